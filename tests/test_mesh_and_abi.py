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
    assert np.array_equal(M.read_reference_files(n, p, with_conductor=True)[2], np.zeros(len(tris)))
    cond = np.arange(len(tris)) % 2
    M.write_reference_files(nodes, tris, n, p, conductor=cond)
    n3, t3, c3 = M.read_reference_files(n, p, with_conductor=True)
    assert np.array_equal(t3, tris) and np.array_equal(c3, cond)
    # the C++ reader of include/rcs_amd.hpp (host only, no library needed) reads the same
    # files: with and without the conductor column
    import subprocess
    src = tmp_path / "read.cpp"
    src.write_text('#include "rcs_amd.hpp"\n#include <iostream>\nint main(int, char** a) {\n'
                   '  std::vector<double> x; std::vector<int> t, c;\n'
                   '  rcs::readMeshFiles(a[1], a[2], x, t, &c);\n'
                   '  long s = 0; for (int v : c) s += v; long st = 0; for (int v : t) st += v;\n'
                   '  std::cout << x.size() / 3 << " " << t.size() / 3 << " " << c.size() << " " << s << " " << st;\n}\n')
    exe = tmp_path / "read"
    r = subprocess.run(["g++", "-std=c++17", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    for conductor in (cond, None):
        M.write_reference_files(nodes, tris, n, p, conductor=conductor)
        out = subprocess.run([str(exe), str(n), str(p)], capture_output=True, text=True, timeout=60).stdout.split()
        want_c = 0 if conductor is None else int(cond.sum())
        assert list(map(int, out)) == [len(nodes), len(tris), len(tris), want_c, int(tris.sum())]


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


def test_bad_arguments_are_errors_before_any_device_work():
    """Every entry point checks its arguments on the host before its first device call, so
    on a machine without a GPU each bad call below fails with its own message, not with
    'no HIP device' (and on a GPU it never reaches a kernel)."""
    import rcs_amd
    L = rcs_amd.lib()
    C = ctypes
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int))
    fp = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))

    def err(rc):
        assert rc < 0
        return L.rcs_last_error().decode()

    z = np.zeros(64)
    zi = np.zeros(16, dtype=np.int32)
    zf = np.zeros(64, dtype=np.float32)
    h = C.c_void_p()
    # dense kernels: sizes, operations, leading dimensions, null buffers
    assert "negative" in err(L.rcs_gemm(0, 0, 0, -1, 2, 2, dp(z), 1, dp(z), 2, dp(z)))
    assert "op must be" in err(L.rcs_gemm(0, 2, 0, 2, 2, 2, dp(z), 2, dp(z), 2, dp(z)))
    assert "leading dimension" in err(L.rcs_gemm(0, 0, 0, 4, 2, 2, dp(z), 3, dp(z), 2, dp(z)))
    assert "leading dimension" in err(L.rcs_gemm(0, 1, 1, 4, 2, 3, dp(z), 2, dp(z), 2, dp(z)))
    assert "A is null" in err(L.rcs_gemm(0, 0, 0, 2, 2, 2, None, 2, dp(z), 2, dp(z)))
    assert "C is null" in err(L.rcs_gemm(0, 0, 0, 2, 2, 2, dp(z), 2, dp(z), 2, None))
    assert "partition out of range" in err(L.rcs_rrqr(0, dp(z), 2, 2, 10, 10, 3, 0.0, 0, dp(z), ip(zi), dp(z)))
    assert "negative" in err(L.rcs_rrqr(0, dp(z), -2, 2, 10, 10, 0, 0.0, 0, dp(z), ip(zi), dp(z)))
    assert "rank is null" in err(L.rcs_rrqr(0, dp(z), 2, 2, 10, 10, 0, 0.0, 0, dp(z), None, dp(z)))
    assert "negative" in err(L.rcs_lu_solve(0, dp(z), -1, dp(z), 1, dp(z)))
    assert "X is null" in err(L.rcs_lu_solve(0, dp(z), 2, dp(z), 1, None))
    # meshes: kind, frequency, sizes, null arrays, node numbers (1-based, as in the reference files)
    nodes = np.eye(3)
    tris = np.array([[1, 2, 3]], dtype=np.int32)
    bad = np.array([[1, 2, 4]], dtype=np.int32)
    assert "kernel kind" in err(L.rcs_mesh_create(dp(nodes), 3, ip(tris), 1, None, 7, 1e9, C.byref(h)))
    assert "frequency" in err(L.rcs_mesh_create(dp(nodes), 3, ip(tris), 1, None, 0, 0.0, C.byref(h)))
    assert "at least 3 nodes" in err(L.rcs_mesh_create(dp(nodes), 2, ip(tris), 1, None, 0, 1e9, C.byref(h)))
    assert "nodes is null" in err(L.rcs_mesh_create(None, 3, ip(tris), 1, None, 0, 1e9, C.byref(h)))
    assert "out is null" in err(L.rcs_mesh_create(dp(nodes), 3, ip(tris), 1, None, 0, 1e9, None))
    assert "outside the node list" in err(L.rcs_mesh_create(dp(nodes), 3, ip(bad), 1, None, 0, 1e9, C.byref(h)))
    # null handles and solver parameters
    assert "mesh is null" in err(L.rcs_entries(None, 0, ip(zi), 1, ip(zi), 1, dp(z)))
    assert "mesh is null" in err(L.rcs_aca(None, 0, 1e-3, ip(zi), 1, ip(zi), 1, ip(zi), ip(zi), ip(zi)))
    assert "mesh is null" in err(L.rcs_rhs(None, dp(z), dp(z), dp(z)))
    assert "mesh is null" in err(L.rcs_mesh_locations(None, dp(z)))
    bb = np.array([-1.0, 1, -1, 1, -1, 1])
    assert "mesh is null" in err(L.rcs_solver_create(None, dp(bb), 64, 6, 6, 6, 0, 0, C.byref(h)))
    assert "variant" in err(L.rcs_solver_create(None, dp(bb), 64, 6, 6, 6, 0, 2, C.byref(h)))
    assert "precision" in err(L.rcs_solver_create(None, dp(bb), 64, 6, 6, 6, 5, 0, C.byref(h)))
    assert "leaf size" in err(L.rcs_solver_create(None, dp(bb), 0, 6, 6, 6, 0, 0, C.byref(h)))
    flat = np.array([0.0, 0, 0, 0, 0, 0])
    assert "bbox is empty" in err(L.rcs_solver_create(None, dp(flat), 64, 6, 6, 6, 0, 0, C.byref(h)))
    swapped = np.array([1.0, -1, -1, 1, -1, 1])
    assert "min <= max" in err(L.rcs_operator_create(None, dp(swapped), 64, 10, C.byref(h)))
    assert "solver is null" in err(L.rcs_solver_factorize(None))
    assert "solver is null" in err(L.rcs_solver_solve(None, dp(z), 1, dp(z), None))
    assert "solver is null" in err(L.rcs_solver_stats(None, ip(zi), dp(z)))
    assert "solver is null" in err(L.rcs_solver_ranks(None, 0, ip(zi), 16))
    assert "solver is null" in err(L.rcs_solver_rank_log(None, fp(zf), 1))
    assert "operator is null" in err(L.rcs_operator_apply(None, dp(z), 1, dp(z)))
    assert "solver is null" in err(L.rcs_solver_refine(None, None, dp(z), 1, dp(z), 1e-6, 10, 10, None, None))
    assert "model is null" in err(L.rcs_mlp_forward(None, fp(zf), 1, fp(zf)))
    assert "model is null" in err(L.rcs_mlp_train(None, fp(zf), fp(zf), 1, 1, 0.01, None))
    assert "out is null" in err(L.rcs_mlp_create(1, None))
    assert "profile tag" in err(L.rcs_profile_get(99, dp(z), dp(z), None))
    assert "burst" in err(L.rcs_set_aca_burst(0))


def test_cpp_header_offers_the_reference_surface():
    """tests/cpp/reference_surface.cpp uses every member the reference's driver and tree use
    on its classes (EFIEMatrix's nodeList / edge2PatchMap / conductorIndices, getCoefficients,
    getJAtCentroids, getRCS_OnePointQuadrature, AIFMM's backSubstitute1/2/3 and getPhi,
    ElectrostaticsMatrix's getRhs / getchargeDensityUsingLU / getCapacitance, RRQR's
    errorInQR ...) with the reference's names; it must compile against include/rcs_amd.hpp."""
    import subprocess
    src = os.path.join(ROOT, "tests", "cpp", "reference_surface.cpp")
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                        src], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr


def test_cpp_header_postprocessing_on_host_only_meshes(fixture78, tmp_path):
    """tests/cpp/host_postprocess.cpp runs the reference driver's mesh and post-processing
    flow through the C++ drop-in classes (host-only meshes, linked against librcs_amd.so, no
    GPU): the RWG midpoints it builds from edge2PatchMap and nodeList, the right-hand side,
    getJAtCentroids and getRCS_OnePointQuadrature, and the ElectrostaticsMatrix's getRhs /
    getCapacitance / getCentroids, each against the oracle."""
    import subprocess
    from oracle import pyoracle as O
    from rcs_amd import mesh as M
    nodes, tris = fixture78
    n, p = tmp_path / "nodeList.txt", tmp_path / "patchList.txt"
    M.write_reference_files(nodes, tris, n, p)
    exe = tmp_path / "host_postprocess"
    libdir = os.path.join(ROOT, "rcs_amd")
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"),
                        os.path.join(ROOT, "tests", "cpp", "host_postprocess.cpp"), "-L", libdir, "-lrcs_amd",
                        f"-Wl,-rpath,{libdir}", "-o", str(exe)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    out = str(tmp_path / "pp")
    r = subprocess.run([str(exe), str(n), str(p), out], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    o = O.Mesh(nodes, tris)
    N, ncond, npatch = map(int, open(out + "_info.txt").read().split())
    assert (N, ncond, npatch) == (o.n_rwg, 1, o.n_patch)
    assert np.allclose(np.loadtxt(out + "_mid.txt"), o.rwg_midpoints(), rtol=0, atol=1e-15)
    cplx = lambda f: (lambda a: a[:, 0] + 1j * a[:, 1])(np.loadtxt(f, ndmin=2))
    rhs = cplx(out + "_rhs.txt")
    ro = o.rhs(1e9)
    assert np.linalg.norm(rhs - ro) < 1e-13 * np.linalg.norm(ro)
    coef = 0.001 * (np.arange(N) + 1) + 1j * (0.5 - 0.0007 * np.arange(N))
    J = cplx(out + "_J.txt").reshape((npatch, 3), order="F")
    Jo = o.j_at_centroids(1e9, coef)
    assert np.linalg.norm(J - Jo) < 1e-13 * np.linalg.norm(Jo)
    th_s = np.loadtxt(out + "_rcs.txt")
    so = o.rcs(1e9, coef, th_s[:, 0], np.zeros(len(th_s)))
    assert np.linalg.norm(th_s[:, 1] - so) < 1e-12 * np.linalg.norm(so)
    lap = open(out + "_lap.txt").read().split("\n")
    nl, ones, cap = lap[0].split()
    assert int(nl) == npatch and float(ones) == npatch  # conductor 0 everywhere: rhs all ones
    pa = o.patches()
    assert abs(float(cap) - pa[:, 3].sum()) < 1e-14 * pa[:, 3].sum()
    assert np.allclose(np.loadtxt(lap[1:]), pa[:, :3], rtol=0, atol=1e-15)


def test_c_header_is_plain_c(tmp_path):
    """include/rcs_amd.h is usable from C (C99, pedantic): the boundary takes plain pointers and
    sizes, so any FFI that reads C declarations can bind it."""
    import subprocess
    src = tmp_path / "c_check.c"
    src.write_text('#include "rcs_amd.h"\nint main(void) { rcs_mesh_t m = 0; rcs_solver_t s = 0; (void)m; (void)s;'
                   ' return 0; }\n')
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Wextra", "-pedantic", "-Werror", "-fsyntax-only", "-I",
                        os.path.join(ROOT, "include"), str(src)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
