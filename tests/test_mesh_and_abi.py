"""Host-side checks: mesh generators, reference file format, and the C ABI surface."""
import ctypes
import os
import re

import numpy as np
import pytest

from rcs_amd import mesh as M

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("level", [0, 1, 2, 3])
def test_icosphere_closed(level):
    nodes, tris = M.icosphere(level, 0.15)
    assert tris.shape[0] == 20 * 4 ** level
    assert np.allclose(np.linalg.norm(nodes, axis=1), 0.15)
    edges = {}
    for t in tris:
        for a, b in ((t[0], t[1]), (t[1], t[2]), (t[2], t[0])):
            k = (min(a, b), max(a, b))
            edges[k] = edges.get(k, 0) + 1
    assert set(edges.values()) == {2}
    assert len(edges) == 3 * tris.shape[0] // 2  # closed surface: N = 3F/2 RWG bases


def test_reference_format_roundtrip(tmp_path):
    nodes, tris = M.ellipsoid(2, (0.1, 0.2, 0.3))
    n, p = tmp_path / "nodeList.txt", tmp_path / "patchList.txt"
    M.write_reference_files(nodes, tris, n, p)
    n2, t2 = M.read_reference_files(n, p)
    assert np.array_equal(t2, tris) and np.allclose(n2, nodes, rtol=0, atol=1e-17)


def test_perturbed_ellipsoid_seeded():
    a = M.perturbed_ellipsoid(2, (0.15, 0.15, 0.2), 0.02, seed=7)[0]
    b = M.perturbed_ellipsoid(2, (0.15, 0.15, 0.2), 0.02, seed=7)[0]
    c = M.perturbed_ellipsoid(2, (0.15, 0.15, 0.2), 0.02, seed=8)[0]
    assert np.array_equal(a, b) and not np.array_equal(a, c)


def _declared_symbols():
    hdr = open(os.path.join(ROOT, "include", "rcs_amd.h")).read()
    return sorted(set(re.findall(r"\b(rcs_[a-z_0-9]+)\s*\(", hdr)))


def test_abi_exports_every_declared_symbol():
    import rcs_amd
    lib = ctypes.CDLL(rcs_amd.LIB_PATH)
    names = _declared_symbols()
    assert len(names) >= 18
    for n in names:
        assert hasattr(lib, n), f"{n} declared in include/rcs_amd.h but not exported"
    assert sorted(rcs_amd.EXPORTED) == names


def test_abi_error_string_without_gpu_work():
    import rcs_amd
    lib = rcs_amd.lib()
    assert isinstance(lib.rcs_last_error(), bytes)


def test_only_lane_zero_is_accepted():
    """Concurrent lanes are switched off: lane 0 is accepted, any other lane is an error with a
    message (no device work is involved)."""
    import rcs_amd
    L = rcs_amd.lib()
    assert L.rcs_set_lane(0) == 0
    assert L.rcs_set_lane(1) < 0
    assert b"lane 0 only" in L.rcs_last_error()
    assert L.rcs_set_lane(64) < 0
