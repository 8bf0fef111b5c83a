"""ORACLE — test infrastructure only.

ctypes view of liboracle.so, the FP64 CPU restatement of the reference path. Used by
tests/ for the expected results, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg.
The product path (rcs_amd/) never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

EFIE, LAPLACE, EFIE_PHYSICAL_PROJECTION = 0, 1, 2


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run make -C oracle)")
        L = C.CDLL(path)
        vp, ip, dp = C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_double)
        L.orc_last_error.restype = C.c_char_p
        L.orc_set_dense_threads.argtypes = [C.c_int]
        L.orc_mesh_create.restype = vp
        L.orc_mesh_create.argtypes = [dp, C.c_int, ip, C.c_int, ip]
        L.orc_mesh_destroy.argtypes = [vp]
        L.orc_mesh_nrwg.argtypes = [vp]
        L.orc_mesh_npatch.argtypes = [vp]
        L.orc_mesh_rwg.argtypes = [vp, ip, dp]
        L.orc_mesh_patches.argtypes = [vp, dp]
        L.orc_entries.argtypes = [vp, C.c_int, C.c_double, ip, C.c_int, ip, C.c_int, dp]
        L.orc_rhs.argtypes = [vp, C.c_double, dp, dp, dp]
        L.orc_rcs.argtypes = [vp, C.c_double, dp, dp, dp, C.c_int, dp]
        L.orc_j_at_centroids.argtypes = [vp, C.c_double, dp, dp]
        L.orc_efie_integrals.argtypes = [vp, C.c_int, C.c_double, dp, C.c_int, C.c_int, dp, dp]
        L.orc_aca.argtypes = [vp, C.c_int, C.c_double, ip, C.c_int, ip, C.c_int, C.c_double, ip, ip, ip]
        L.orc_aca_dense.argtypes = [dp, C.c_int, C.c_int, C.c_double, ip, ip, ip]
        L.orc_map_walk_cost.argtypes = [vp, C.c_double, C.c_int, C.c_uint, dp, dp]
        L.orc_rrqr.argtypes = [dp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int,
                               dp, dp, ip, ip, dp]
        L.orc_lu_solve.argtypes = [dp, C.c_int, dp, C.c_int, dp]
        L.orc_dense_solve.argtypes = [vp, C.c_int, C.c_double, dp, dp]
        L.orc_matvec.argtypes = [vp, C.c_int, C.c_double, dp, dp]
        L.orc_aifmm_create.restype = vp
        L.orc_aifmm_create.argtypes = [vp, C.c_int, C.c_double, dp, C.c_int, dp, C.c_int, C.c_int, C.c_int,
                                       C.c_int, C.c_int]
        L.orc_aifmm_destroy.argtypes = [vp]
        L.orc_aifmm_factorize.argtypes = [vp]
        L.orc_aifmm_solve.argtypes = [vp, dp, dp, dp]
        L.orc_aifmm_info.argtypes = [vp, ip, dp, dp]
        L.orc_aifmm_ranks.argtypes = [vp, C.c_int, ip]
        L.orc_aifmm_tree_summary.argtypes = [vp, dp]
        _LIB = L
    return _LIB


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _check(rc):
    if rc != 0:
        raise RuntimeError("oracle: " + lib().orc_last_error().decode())


class Mesh:
    """RWG mesh (src/EFIEMatrix.cpp:5-209) built from node/triangle arrays."""

    def __init__(self, nodes, tris, conductor=None):
        nodes = np.ascontiguousarray(nodes, dtype=np.float64)
        tris = np.ascontiguousarray(tris, dtype=np.int32)
        cond = None if conductor is None else np.ascontiguousarray(conductor, dtype=np.int32)
        self.h = lib().orc_mesh_create(_dp(nodes), nodes.shape[0], _ip(tris), tris.shape[0],
                                       None if cond is None else _ip(cond))
        self.nodes = nodes
        self.tris = tris
        self.n_rwg = lib().orc_mesh_nrwg(self.h)
        self.n_patch = lib().orc_mesh_npatch(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_mesh_destroy(self.h)
            self.h = None

    def rwg(self):
        t = np.zeros((self.n_rwg, 6), dtype=np.int32)
        ln = np.zeros(self.n_rwg)
        lib().orc_mesh_rwg(self.h, _ip(t), _dp(ln))
        return t, ln

    def patches(self):
        p = np.zeros((self.n_patch, 5))
        lib().orc_mesh_patches(self.h, _dp(p))
        return p

    def rwg_midpoints(self):
        """Unknown locations of the EFIE solver: RWG edge midpoints (S:143-152)."""
        t, _ = self.rwg()
        full = np.vstack([np.zeros((1, 3)), self.nodes])
        return (full[t[:, 0]] + full[t[:, 1]]) / 2.0

    def centroids(self):
        return self.patches()[:, :3].copy()

    def entries(self, kind, freq, rows, cols):
        rows = np.ascontiguousarray(rows, dtype=np.int32)
        cols = np.ascontiguousarray(cols, dtype=np.int32)
        out = np.zeros(rows.size * cols.size, dtype=np.complex128)
        _check(lib().orc_entries(self.h, kind, freq, _ip(rows), rows.size, _ip(cols), cols.size,
                                 out.ctypes.data_as(C.POINTER(C.c_double))))
        return out.reshape(cols.size, rows.size).T

    def rhs(self, freq, E=(1.0, 0.0, 0.0), kvec=(0.0, 0.0, 1.0)):
        out = np.zeros(self.n_rwg, dtype=np.complex128)
        E = np.asarray(E, dtype=np.float64)
        k = np.asarray(kvec, dtype=np.float64)
        _check(lib().orc_rhs(self.h, freq, _dp(E), _dp(k), out.ctypes.data_as(C.POINTER(C.c_double))))
        return out

    def rcs(self, freq, coef, theta, phi):
        coef = np.ascontiguousarray(coef, dtype=np.complex128)
        theta = np.ascontiguousarray(theta, dtype=np.float64)
        phi = np.ascontiguousarray(phi, dtype=np.float64)
        out = np.zeros(theta.size)
        _check(lib().orc_rcs(self.h, freq, coef.ctypes.data_as(C.POINTER(C.c_double)), _dp(theta), _dp(phi),
                             theta.size, _dp(out)))
        return out

    def efie_integrals(self, kind, freq, tc, tri, fv):
        """(scalar, vector) Green's-function integrals over patch tri seen from tc, the vector
        one weighted by r' - node fv (E:670-685, E:386-446)."""
        tc = np.ascontiguousarray(tc, dtype=np.float64)
        s = np.zeros(2)
        v = np.zeros(6)
        _check(lib().orc_efie_integrals(self.h, kind, freq, _dp(tc), tri, fv, _dp(s), _dp(v)))
        return s[0] + 1j * s[1], v[0::2] + 1j * v[1::2]

    def j_at_centroids(self, freq, coef):
        coef = np.ascontiguousarray(coef, dtype=np.complex128)
        J = np.zeros((self.n_patch, 3), dtype=np.complex128, order="F")
        _check(lib().orc_j_at_centroids(self.h, freq, coef.ctypes.data_as(C.POINTER(C.c_double)),
                                        J.ctypes.data_as(C.POINTER(C.c_double))))
        return J

    def aca(self, kind, freq, rows, cols, tol):
        rows = np.ascontiguousarray(rows, dtype=np.int32)
        cols = np.ascontiguousarray(cols, dtype=np.int32)
        m = min(rows.size, cols.size)
        rk = np.zeros(1, dtype=np.int32)
        rb = np.zeros(m, dtype=np.int32)
        cb = np.zeros(m, dtype=np.int32)
        _check(lib().orc_aca(self.h, kind, freq, _ip(rows), rows.size, _ip(cols), cols.size, tol,
                             _ip(rk), _ip(rb), _ip(cb)))
        r = int(rk[0])
        return r, rb[:r].copy(), cb[:r].copy()

    def map_walk_cost(self, freq, samples=2000, seed=1):
        """Seconds per entry of the reference's std::map basis lookup and of the restatement's
        whole EFIE entry, timed on the same random (i, j) pairs on one thread."""
        w = np.zeros(1)
        e = np.zeros(1)
        _check(lib().orc_map_walk_cost(self.h, freq, samples, seed, _dp(w), _dp(e)))
        return float(w[0]), float(e[0])

    def dense_solve(self, kind, freq, rhs):
        rhs = np.ascontiguousarray(rhs, dtype=np.complex128)
        x = np.zeros_like(rhs)
        _check(lib().orc_dense_solve(self.h, kind, freq, rhs.ctypes.data_as(C.POINTER(C.c_double)),
                                     x.ctypes.data_as(C.POINTER(C.c_double))))
        return x

    def matvec(self, kind, freq, x):
        x = np.ascontiguousarray(x, dtype=np.complex128)
        y = np.zeros_like(x)
        _check(lib().orc_matvec(self.h, kind, freq, x.ctypes.data_as(C.POINTER(C.c_double)),
                                y.ctypes.data_as(C.POINTER(C.c_double))))
        return y


def set_dense_threads(n):
    """Threads of the dense products, LU and solves (1, the reference driver's policy, by
    default); returns the previous count."""
    return int(lib().orc_set_dense_threads(int(n)))


def aca_dense(A, tol):
    """ACA_only_nodes2 over every row and column of the complex matrix A: (rank, pivot rows,
    pivot columns)."""
    A = np.asfortranarray(A, dtype=np.complex128)
    m, n = A.shape
    mn = min(m, n)
    rk = np.zeros(1, dtype=np.int32)
    rb = np.zeros(mn, dtype=np.int32)
    cb = np.zeros(mn, dtype=np.int32)
    _check(lib().orc_aca_dense(A.ctypes.data_as(C.POINTER(C.c_double)), m, n, tol, _ip(rk), _ip(rb), _ip(cb)))
    r = int(rk[0])
    return r, rb[:r].copy(), cb[:r].copy()


def rrqr(A, tol_pow1, tol_pow2, partition, input_pivot, fillin_rank):
    A = np.asfortranarray(A, dtype=np.complex128)
    m, n = A.shape
    mn = min(m, n)
    Q = np.zeros((m, mn), dtype=np.complex128, order="F")
    R = np.zeros((mn, n), dtype=np.complex128, order="F")
    perm = np.zeros(n, dtype=np.int32)
    rk = np.zeros(1, dtype=np.int32)
    mp = np.zeros(1)
    _check(lib().orc_rrqr(A.ctypes.data_as(C.POINTER(C.c_double)), m, n, tol_pow1, tol_pow2, partition,
                          input_pivot, fillin_rank, Q.ctypes.data_as(C.POINTER(C.c_double)),
                          R.ctypes.data_as(C.POINTER(C.c_double)), _ip(perm), _ip(rk), _dp(mp)))
    r = int(rk[0])
    Qf = Q.reshape(-1, order="F")[: m * r].reshape((m, r), order="F")
    Rf = R.reshape(-1, order="F")[: r * n].reshape((r, n), order="F")
    return Qf, Rf, perm, r, float(mp[0])


def lu_solve(A, B):
    A = np.asfortranarray(A, dtype=np.complex128)
    B = np.asfortranarray(B.reshape(A.shape[0], -1), dtype=np.complex128)
    X = np.zeros_like(B, order="F")
    _check(lib().orc_lu_solve(A.ctypes.data_as(C.POINTER(C.c_double)), A.shape[0],
                              B.ctypes.data_as(C.POINTER(C.c_double)), B.shape[1],
                              X.ctypes.data_as(C.POINTER(C.c_double))))
    return X


class AIFMM:
    """AIFMM<K> (src/AIFMM.cpp): construction = assembly, then factorize/solve."""

    def __init__(self, mesh: Mesh, kind, freq, locations, bbox, leaf_size, tol_pow, fillin_tol,
                 fillin_rank=0, variant=0):
        loc = np.ascontiguousarray(locations, dtype=np.float64)
        bbox = np.ascontiguousarray(bbox, dtype=np.float64)
        self.mesh = mesh
        self.N = loc.shape[0]
        self.h = lib().orc_aifmm_create(mesh.h, kind, freq, _dp(loc), self.N, _dp(bbox), leaf_size, tol_pow,
                                        fillin_tol, fillin_rank, variant)
        if not self.h:
            raise RuntimeError("oracle: " + lib().orc_last_error().decode())

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_aifmm_destroy(self.h)
            self.h = None

    def factorize(self):
        _check(lib().orc_aifmm_factorize(self.h))

    def solve(self, rhs, box_order=False):
        rhs = np.ascontiguousarray(rhs, dtype=np.complex128)
        x = np.zeros_like(rhs)
        xb = np.zeros_like(rhs)
        _check(lib().orc_aifmm_solve(self.h, rhs.ctypes.data_as(C.POINTER(C.c_double)),
                                     x.ctypes.data_as(C.POINTER(C.c_double)),
                                     xb.ctypes.data_as(C.POINTER(C.c_double))))
        return (x, xb) if box_order else x

    def info(self):
        i = np.zeros(5, dtype=np.int32)
        t = np.zeros(3)
        f = np.zeros(4)
        lib().orc_aifmm_info(self.h, _ip(i), _dp(t), _dp(f))
        return {"nLevels": int(i[0]), "startLevel": int(i[1]), "n_rrqr": int(i[2]),
                "n_compress_p2p": int(i[3]), "n_compress_p2l": int(i[4]),
                "t_assemble": t[0], "t_factor": t[1], "t_solve": t[2],
                "gemm_flops": f[0], "getrf_flops": f[1], "getrs_flops": f[2], "qr_flops": f[3]}

    def ranks(self, level):
        out = np.zeros(3 * 8 ** level, dtype=np.int32)
        n = lib().orc_aifmm_ranks(self.h, level, _ip(out))
        return out[: 3 * n].reshape(n, 3)

    def tree_summary(self):
        """getAvgLeafSize, getAvgMultipoles, findMemory, findMemoryAfterFactorize (GB)."""
        out = np.zeros(4)
        lib().orc_aifmm_tree_summary(self.h, _dp(out))
        return {"avg_leaf_size": int(out[0]), "avg_multipoles": int(out[1]), "memory_gb": out[2],
                "memory_after_factorize_gb": out[3]}


def panel(tc, a, b, c):
    """Closed-form integrals of 1/R and (r'-p)/R over triangle (a, b, c) from tc (p: projection of tc)."""
    L = lib()
    dp = C.POINTER(C.c_double)
    L.orc_panel.argtypes = [dp, dp, dp, dp, dp, dp]
    arr = [np.ascontiguousarray(x, dtype=np.float64) for x in (tc, a, b, c)]
    lap = np.zeros(1)
    rho = np.zeros(3)
    L.orc_panel(*[_dp(x) for x in arr], _dp(lap), _dp(rho))
    return float(lap[0]), rho


def aifmm_block(solver, level, box, which, key=0, cap=4_000_000):
    """One stored block of the oracle tree (see rcs_solver_block for `which`)."""
    L = lib()
    L.orc_aifmm_block.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int),
                                  C.POINTER(C.c_double), C.c_longlong]
    shape = np.zeros(2, dtype=np.int32)
    out = np.zeros(cap, dtype=np.complex128)
    rc = L.orc_aifmm_block(solver.h, level, box, which, key, _ip(shape), out.ctypes.data_as(C.POINTER(C.c_double)),
                           cap)
    if rc != 0:
        return None
    r, c = int(shape[0]), int(shape[1])
    return out[: r * c].reshape((r, c), order="F").copy()
