"""Host-side arithmetic of bench.py (no GPU): the solve's roofline figures and the refusal of
more than one GPU."""
import os
import subprocess
import sys

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_solve_roofline_by_hand():
    # 1 GB of factors, N = 1000, complex128: one RHS moves 1e9 + 2*1000*16 bytes in 0.5 s
    r = bench.solve_roofline(1.0e9, 1000, 0, 0.5, 2.0)
    assert r["bytes_1rhs"] == 1.0e9 + 32000
    assert abs(r["achieved_1rhs"] - (1.0e9 + 32000) / 0.5 / 1e9) < 1e-12
    assert abs(r["frac_1rhs"] - (1.0e9 + 32000) / 0.5 / 8.0e12) < 1e-15
    assert r["bytes_181rhs"] == 1.0e9 + 2 * 1000 * 181 * 16
    # 1e9 / 16 complex128 entries, 8 flop each per column, 181 columns, 2 s
    assert abs(r["achieved_tflops_181rhs"] - 8 * 1.0e9 / 16 * 181 / 2.0 / 1e12) < 1e-12
    assert abs(r["frac_flops_181rhs"] - r["achieved_tflops_181rhs"] * 1e12 / 78.6e12) < 1e-15
    # complex64 halves the vector bytes and uses the FP32 peak
    r = bench.solve_roofline(1.0e9, 1000, 1, 0.5, 2.0)
    assert r["bytes_1rhs"] == 1.0e9 + 16000
    assert abs(r["frac_flops_181rhs"] - r["achieved_tflops_181rhs"] * 1e12 / 157.3e12) < 1e-15
    # a zero time reports zero, not a division error
    assert bench.solve_roofline(1.0e9, 1000, 0, 0.0, 0.0)["achieved_1rhs"] == 0.0


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "exactly one GPU" in r.stderr


def test_extra_fp64_forms():
    import argparse

    class FakeR:
        C128 = 0
    a = argparse.Namespace(extra_fp64_tols="3, 2,4", mixed_tol=4)
    assert bench.extra_fp64_forms(FakeR, a) == [("fp64_tol3_refined", 0, 3, False), ("fp64_tol2_refined", 0, 2, False)]
    assert bench.extra_fp64_forms(FakeR, argparse.Namespace(extra_fp64_tols="", mixed_tol=4)) == []


def test_reference_map_walk_reports_both_costs():
    r = bench.reference_map_walk(2, 0.15, 1e9, samples=200)
    assert r["N"] == 3 * 20 * 4 ** 2 // 2 and r["samples"] == 200
    assert r["lookup_us_per_entry"] > 0 and r["port_entry_us"] > 0
    assert abs(r["lookup_over_entry"] - r["lookup_us_per_entry"] / r["port_entry_us"]) < 1e-9
