"""MI355X-native AIFMM fast direct solver (EFIE / Laplace single layer).

Python view of the C ABI in include/rcs_amd.h (librcs_amd.so, hand-written HIP for
gfx950). The classes mirror the reference's C++ interface: ``Mesh`` stands for
EFIEMatrix / ElectrostaticsMatrix (mesh ingest, getMatrixEntry, getRhs, RCS),
``AIFMM`` for AIFMM<K> (construction = assembly, factorize, solve), plus the ACA,
RRQR, LU and GEMM building blocks. Nothing here falls back to a CPU path: a missing
or unloadable library raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import mesh  # noqa: F401

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librcs_amd.so")
_LIB = None

EFIE, LAPLACE = 0, 1
C128, C64 = 0, 1


def lib():
    """Load librcs_amd.so (raises if it is not built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        vp, ip, dp = C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_double)
        L.rcs_last_error.restype = C.c_char_p
        L.rcs_mesh_create.argtypes = [dp, C.c_int, ip, C.c_int, ip, C.c_int, C.c_double, C.POINTER(vp)]
        L.rcs_mesh_create_host.argtypes = [dp, C.c_int, ip, C.c_int, ip, C.c_int, C.c_double, C.POINTER(vp)]
        L.rcs_mesh_npatch.argtypes = [vp]
        L.rcs_mesh_patches.argtypes = [vp, dp, ip]
        L.rcs_j_at_centroids.argtypes = [vp, dp, dp]
        L.rcs_rcs_from_j.argtypes = [vp, dp, dp, dp, C.c_int, dp]
        L.rcs_mesh_destroy.argtypes = [vp]
        L.rcs_mesh_size.argtypes = [vp]
        L.rcs_mesh_locations.argtypes = [vp, dp]
        L.rcs_mesh_rwg.argtypes = [vp, ip, dp]
        L.rcs_entries.argtypes = [vp, C.c_int, ip, C.c_int, ip, C.c_int, dp]
        L.rcs_rhs.argtypes = [vp, dp, dp, dp]
        L.rcs_rcs.argtypes = [vp, dp, dp, dp, C.c_int, dp]
        L.rcs_aca.argtypes = [vp, C.c_int, C.c_double, ip, C.c_int, ip, C.c_int, ip, ip, ip]
        L.rcs_rrqr.argtypes = [C.c_int, dp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int,
                               dp, ip, dp]
        L.rcs_lu_solve.argtypes = [C.c_int, dp, C.c_int, dp, C.c_int, dp]
        L.rcs_gemm.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, dp, C.c_int, dp, C.c_int, dp]
        L.rcs_solver_create.argtypes = [vp, dp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                        C.POINTER(vp)]
        L.rcs_solver_destroy.argtypes = [vp]
        L.rcs_solver_factorize.argtypes = [vp]
        L.rcs_solver_solve.argtypes = [vp, dp, C.c_int, dp, dp]
        L.rcs_solver_stats.argtypes = [vp, ip, dp]
        L.rcs_solver_ranks.argtypes = [vp, C.c_int, ip, C.c_int]
        L.rcs_solver_tree_summary.argtypes = [vp, dp]
        L.rcs_profile_enable.argtypes = [C.c_int]
        L.rcs_profile_get.argtypes = [C.c_int, dp, dp, C.POINTER(C.c_longlong)]
        L.rcs_operator_create.argtypes = [vp, dp, C.c_int, C.c_int, C.POINTER(vp)]
        L.rcs_operator_destroy.argtypes = [vp]
        L.rcs_operator_apply.argtypes = [vp, dp, C.c_int, dp]
        L.rcs_solver_refine.argtypes = [vp, vp, dp, C.c_int, dp, C.c_double, C.c_int, C.c_int, ip, dp]
        fp = C.POINTER(C.c_float)
        L.rcs_mlp_create.argtypes = [C.c_uint, C.POINTER(vp)]
        L.rcs_mlp_destroy.argtypes = [vp]
        L.rcs_mlp_get.argtypes = [vp, fp, fp, fp]
        L.rcs_mlp_set.argtypes = [vp, fp, fp, fp]
        L.rcs_mlp_forward.argtypes = [vp, fp, C.c_int, fp]
        L.rcs_mlp_gradient.argtypes = [vp, fp, fp, C.c_int, fp, dp]
        L.rcs_mlp_train.argtypes = [vp, fp, fp, C.c_int, C.c_int, C.c_float, dp]
        L.rcs_set_rank_model.argtypes = [vp]
        L.rcs_set_aca_burst.argtypes = [C.c_int]
        L.rcs_set_lane.argtypes = [C.c_int]
        L.rcs_solver_rank_log.argtypes = [vp, fp, C.c_int]
        L.rcs_solver_aca_stats.argtypes = [vp, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong), dp]
        _LIB = L
    return _LIB


EXPORTED = [
    "rcs_init", "rcs_last_error", "rcs_mesh_create", "rcs_mesh_create_host", "rcs_mesh_destroy", "rcs_mesh_size",
    "rcs_mesh_locations", "rcs_mesh_rwg", "rcs_mesh_npatch", "rcs_mesh_patches", "rcs_entries", "rcs_rhs", "rcs_rcs",
    "rcs_j_at_centroids", "rcs_rcs_from_j", "rcs_aca", "rcs_rrqr", "rcs_lu_solve", "rcs_gemm",
    "rcs_solver_create", "rcs_solver_destroy", "rcs_solver_factorize", "rcs_solver_solve", "rcs_solver_stats",
    "rcs_solver_ranks", "rcs_solver_tree_summary", "rcs_solver_block", "rcs_profile_enable", "rcs_profile_reset", "rcs_profile_get",
    "rcs_operator_create", "rcs_operator_destroy", "rcs_operator_apply", "rcs_solver_refine",
    "rcs_mlp_create", "rcs_mlp_destroy", "rcs_mlp_get", "rcs_mlp_set", "rcs_mlp_forward", "rcs_mlp_gradient",
    "rcs_mlp_train", "rcs_set_rank_model", "rcs_set_aca_burst", "rcs_set_lane", "rcs_solver_rank_log", "rcs_solver_aca_stats",
]

PROFILE_TAGS = ["schur", "getrf", "getrs", "entries", "aca", "rrqr", "update"]


def set_profiling(on=True):
    _check(lib().rcs_profile_enable(1 if on else 0))


def reset_profile():
    _check(lib().rcs_profile_reset())


def profile():
    """Event-timed totals per phase: {tag: {"flops"|"work", "ms", "launches"}}."""
    out = {}
    for i, t in enumerate(PROFILE_TAGS):
        w, ms = C.c_double(), C.c_double()
        n = C.c_longlong()
        _check(lib().rcs_profile_get(i, C.byref(w), C.byref(ms), C.byref(n)))
        out[t] = {"flops": w.value, "ms": ms.value, "launches": n.value}
    return out


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _check(rc):
    if rc != 0:
        raise RuntimeError("rcs_amd: " + lib().rcs_last_error().decode())


class Mesh:
    """EFIEMatrix / ElectrostaticsMatrix over the GPU geometry tables. With host_only=True
    the mesh keeps host tables only (no device is touched): numbering, patches, right-hand
    sides and RCS work, entries, ACA and solvers do not."""

    def __init__(self, nodes, tris, kind=EFIE, freq=1.0e9, conductor=None, host_only=False):
        nodes = np.ascontiguousarray(nodes, dtype=np.float64)
        tris = np.ascontiguousarray(tris, dtype=np.int32)
        cond = None if conductor is None else np.ascontiguousarray(conductor, dtype=np.int32)
        h = C.c_void_p()
        create = lib().rcs_mesh_create_host if host_only else lib().rcs_mesh_create
        _check(create(_dp(nodes), nodes.shape[0], _ip(tris), tris.shape[0], None if cond is None else _ip(cond),
                      kind, freq, C.byref(h)))
        self.h = h
        self.n_patch = lib().rcs_mesh_npatch(h)
        self.kind = kind
        self.freq = freq
        self.N = lib().rcs_mesh_size(h)

    def __del__(self):
        if getattr(self, "h", None) and _LIB is not None:
            _LIB.rcs_mesh_destroy(self.h)
            self.h = None

    def locations(self):
        out = np.zeros((self.N, 3))
        _check(lib().rcs_mesh_locations(self.h, _dp(out)))
        return out

    def rwg(self):
        t = np.zeros((self.N, 6), dtype=np.int32)
        ln = np.zeros(self.N)
        _check(lib().rcs_mesh_rwg(self.h, _ip(t), _dp(ln)))
        return t, ln

    def entries(self, rows, cols, precision=C128):
        rows = np.ascontiguousarray(rows, dtype=np.int32)
        cols = np.ascontiguousarray(cols, dtype=np.int32)
        out = np.zeros(rows.size * cols.size, dtype=np.complex128)
        _check(lib().rcs_entries(self.h, precision, _ip(rows), rows.size, _ip(cols), cols.size,
                                 out.view(np.float64).ctypes.data_as(C.POINTER(C.c_double))))
        return out.reshape(cols.size, rows.size).T

    def rhs(self, E=(1.0, 0.0, 0.0), kvec=(0.0, 0.0, 1.0)):
        E = np.asarray(E, dtype=np.float64)
        k = np.asarray(kvec, dtype=np.float64)
        out = np.zeros(self.N, dtype=np.complex128)
        _check(lib().rcs_rhs(self.h, _dp(E), _dp(k), out.view(np.float64).ctypes.data_as(C.POINTER(C.c_double))))
        return out

    def rcs(self, coef, theta, phi):
        coef = np.ascontiguousarray(coef, dtype=np.complex128)
        theta = np.ascontiguousarray(theta, dtype=np.float64)
        phi = np.ascontiguousarray(phi, dtype=np.float64)
        out = np.zeros(theta.size)
        _check(lib().rcs_rcs(self.h, coef.view(np.float64).ctypes.data_as(C.POINTER(C.c_double)), _dp(theta),
                             _dp(phi), theta.size, _dp(out)))
        return out

    def patches(self):
        """(centroid xyz, area, largestEdge) per patch and the conductor indices."""
        out = np.zeros((self.n_patch, 5))
        cond = np.zeros(self.n_patch, dtype=np.int32)
        _check(lib().rcs_mesh_patches(self.h, _dp(out), _ip(cond)))
        return out, cond

    def j_at_centroids(self, coef):
        """getJAtCentroids: n_patch x 3 complex surface current."""
        coef = np.ascontiguousarray(coef, dtype=np.complex128)
        J = np.zeros((self.n_patch, 3), dtype=np.complex128, order="F")
        _check(lib().rcs_j_at_centroids(self.h, _cplx_ptr(coef), _cplx_ptr_f(J)))
        return J

    def rcs_from_j(self, J, theta, phi):
        """getRCS_OnePointQuadrature at each (theta, phi)."""
        J = np.asfortranarray(J, dtype=np.complex128)
        theta = np.ascontiguousarray(theta, dtype=np.float64)
        phi = np.ascontiguousarray(phi, dtype=np.float64)
        out = np.zeros(theta.size)
        _check(lib().rcs_rcs_from_j(self.h, _cplx_ptr_f(J), _dp(theta), _dp(phi), theta.size, _dp(out)))
        return out

    def aca(self, rows, cols, tol, precision=C128):
        rows = np.ascontiguousarray(rows, dtype=np.int32)
        cols = np.ascontiguousarray(cols, dtype=np.int32)
        m = max(1, min(rows.size, cols.size))
        rk = np.zeros(1, dtype=np.int32)
        rb = np.zeros(m, dtype=np.int32)
        cb = np.zeros(m, dtype=np.int32)
        _check(lib().rcs_aca(self.h, precision, tol, _ip(rows), rows.size, _ip(cols), cols.size, _ip(rk), _ip(rb),
                             _ip(cb)))
        r = int(rk[0])
        return r, rb[:r].copy(), cb[:r].copy()


def _cplx_ptr(a):
    return a.view(np.float64).ctypes.data_as(C.POINTER(C.c_double))


def _cplx_ptr_f(a):
    """Pointer to a Fortran-ordered complex array's memory (column-major, interleaved)."""
    assert a.flags.f_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


def rrqr(A, tol_pow1, tol_pow2, partition, input_pivot, fillin_rank, precision=C128):
    A = np.asfortranarray(A, dtype=np.complex128)
    m, n = A.shape
    mn = max(1, min(m, n))
    Q = np.zeros(m * mn, dtype=np.complex128)
    rk = np.zeros(1, dtype=np.int32)
    mp = np.zeros(1)
    Af = A.reshape(-1, order="F").copy()
    _check(lib().rcs_rrqr(precision, _cplx_ptr(Af), m, n, tol_pow1, tol_pow2, partition,
                          input_pivot, fillin_rank, _cplx_ptr(Q), _ip(rk), _dp(mp)))
    r = int(rk[0])
    return Q[: m * r].reshape((m, r), order="F"), r, float(mp[0])


def lu_solve(A, B, precision=C128):
    A = np.asarray(A, dtype=np.complex128)
    n = A.shape[0]
    B2 = np.asarray(B, dtype=np.complex128).reshape(n, -1)
    Af = A.reshape(-1, order="F").copy()
    Bf = B2.reshape(-1, order="F").copy()
    X = np.zeros_like(Bf)
    _check(lib().rcs_lu_solve(precision, _cplx_ptr(Af), n, _cplx_ptr(Bf), B2.shape[1], _cplx_ptr(X)))
    return X.reshape(B2.shape, order="F")


def gemm(A, B, opA=0, opB=0, precision=C128):
    A = np.asarray(A, dtype=np.complex128)
    B = np.asarray(B, dtype=np.complex128)
    m = A.shape[0] if opA == 0 else A.shape[1]
    k = A.shape[1] if opA == 0 else A.shape[0]
    n = B.shape[1] if opB == 0 else B.shape[0]
    Af = A.reshape(-1, order="F").copy()
    Bf = B.reshape(-1, order="F").copy()
    Cf = np.zeros(m * n, dtype=np.complex128)
    _check(lib().rcs_gemm(precision, opA, opB, m, n, k, _cplx_ptr(Af), A.shape[0], _cplx_ptr(Bf), B.shape[0],
                          _cplx_ptr(Cf)))
    return Cf.reshape((m, n), order="F")


class AIFMM:
    """AIFMM<K>: construction assembles the compressed operator on the GPU."""

    def __init__(self, mesh_: Mesh, bbox, leaf_size=64, tol_pow=6, fillin_tol=6, fillin_rank=0, precision=C128,
                 variant=0):
        bbox = np.ascontiguousarray(bbox, dtype=np.float64)
        h = C.c_void_p()
        _check(lib().rcs_solver_create(mesh_.h, _dp(bbox), leaf_size, tol_pow, fillin_tol, fillin_rank, precision,
                                       variant, C.byref(h)))
        self.h = h
        self.mesh = mesh_
        self.N = mesh_.N
        self.precision = precision

    def __del__(self):
        if getattr(self, "h", None) and _LIB is not None:
            _LIB.rcs_solver_destroy(self.h)
            self.h = None

    def factorize(self):
        _check(lib().rcs_solver_factorize(self.h))

    def solve(self, rhs, box_order=False):
        rhs = np.asarray(rhs, dtype=np.complex128)
        single = rhs.ndim == 1
        R = rhs.reshape(self.N, -1)
        nrhs = R.shape[1]
        Rf = R.reshape(-1, order="F").copy()
        X = np.zeros_like(Rf)
        XB = np.zeros_like(Rf)
        _check(lib().rcs_solver_solve(self.h, _cplx_ptr(Rf), nrhs, _cplx_ptr(X), _cplx_ptr(XB)))
        X = X.reshape((self.N, nrhs), order="F")
        XB = XB.reshape((self.N, nrhs), order="F")
        if single:
            X, XB = X[:, 0], XB[:, 0]
        return (X, XB) if box_order else X

    def stats(self):
        i = np.zeros(5, dtype=np.int32)
        v = np.zeros(12)
        _check(lib().rcs_solver_stats(self.h, _ip(i), _dp(v)))
        keys = ["t_assemble", "t_factor", "t_solve", "gemm_flops", "getrf_flops", "getrs_flops", "qr_flops",
                "t_getrf", "t_schur", "t_compress", "factor_bytes", "peak_bytes"]
        d = {"nLevels": int(i[0]), "startLevel": int(i[1]), "n_rrqr": int(i[2]), "n_compress_p2p": int(i[3]),
             "n_compress_p2l": int(i[4])}
        d.update({k: float(x) for k, x in zip(keys, v)})
        return d

    def tree_summary(self):
        """getAvgLeafSize, getAvgMultipoles, findMemory, findMemoryAfterFactorize (GB) of the
        current state, as the reference driver prints them."""
        out = np.zeros(4)
        _check(lib().rcs_solver_tree_summary(self.h, _dp(out)))
        return {"avg_leaf_size": int(out[0]), "avg_multipoles": int(out[1]), "memory_gb": out[2],
                "memory_after_factorize_gb": out[3]}

    def ranks(self, level):
        cap = 3 * 8 ** level
        out = np.zeros(cap, dtype=np.int32)
        n = lib().rcs_solver_ranks(self.h, level, _ip(out), cap)
        if n < 0:
            _check(n)
        return out[: 3 * n].reshape(n, 3)

    def refine(self, op, rhs, tol=1e-6, maxit=50, restart=30):
        """GMRES on Z x = rhs with the FP64 operator op, preconditioned by this
        factorization. Returns (x, iterations, final relative residual)."""
        rhs = np.asarray(rhs, dtype=np.complex128)
        single = rhs.ndim == 1
        R = rhs.reshape(self.N, -1)
        nrhs = R.shape[1]
        Rf = R.reshape(-1, order="F").copy()
        X = np.zeros_like(Rf)
        it = np.zeros(1, dtype=np.int32)
        rr = np.zeros(1)
        _check(lib().rcs_solver_refine(self.h, op.h, _cplx_ptr(Rf), nrhs, _cplx_ptr(X), tol, maxit, restart,
                                       _ip(it), _dp(rr)))
        X = X.reshape((self.N, nrhs), order="F")
        return (X[:, 0] if single else X), int(it[0]), float(rr[0])


    def rank_log(self):
        """One row per ACA problem of the assembly: 8 predictor features, rank found,
        rank predicted (0 without a model), level, 0 incoming / 1 outgoing, short side
        (box points), long side (interaction-list points)."""
        n = lib().rcs_solver_rank_log(self.h, None, 0)
        if n < 0:
            _check(n)
        out = np.zeros((max(n, 1), 14), dtype=np.float32)
        lib().rcs_solver_rank_log(self.h, out.ctypes.data_as(C.POINTER(C.c_float)), n)
        return out[:n]

    def aca_stats(self):
        st, sy, sec = C.c_longlong(), C.c_longlong(), C.c_double()
        _check(lib().rcs_solver_aca_stats(self.h, C.byref(st), C.byref(sy), C.byref(sec)))
        return {"steps": st.value, "syncs": sy.value, "seconds": sec.value}


MLP_NP = 4801  # W1[64][8], b1[64], W2[64][64], b2[64], W3[64], b3


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class RankMLP:
    """Learned ACA rank predictor (HIP forward/backward/Adam kernels, BF16 weights)."""

    def __init__(self, seed=1):
        h = C.c_void_p()
        _check(lib().rcs_mlp_create(seed, C.byref(h)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None) and _LIB is not None:
            _LIB.rcs_mlp_destroy(self.h)
            self.h = None

    def get(self):
        p = np.zeros(MLP_NP, dtype=np.float32)
        m = np.zeros(8, dtype=np.float32)
        s = np.zeros(8, dtype=np.float32)
        _check(lib().rcs_mlp_get(self.h, _fp(p), _fp(m), _fp(s)))
        return p, m, s

    def set(self, params, mean, scale):
        p, m, s = (np.ascontiguousarray(a, dtype=np.float32) for a in (params, mean, scale))
        _check(lib().rcs_mlp_set(self.h, _fp(p), _fp(m), _fp(s)))

    def forward(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, 8)
        y = np.zeros(len(X), dtype=np.float32)
        _check(lib().rcs_mlp_forward(self.h, _fp(X), len(X), _fp(y)))
        return y

    def gradient(self, X, t):
        X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, 8)
        t = np.ascontiguousarray(t, dtype=np.float32)
        g = np.zeros(MLP_NP, dtype=np.float32)
        loss = C.c_double()
        _check(lib().rcs_mlp_gradient(self.h, _fp(X), _fp(t), len(X), _fp(g), C.byref(loss)))
        return loss.value, g

    def train(self, X, t, iters=2000, lr=3e-3):
        X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, 8)
        t = np.ascontiguousarray(t, dtype=np.float32)
        loss = C.c_double()
        _check(lib().rcs_mlp_train(self.h, _fp(X), _fp(t), len(X), iters, lr, C.byref(loss)))
        return loss.value

    def predict_ranks(self, X):
        return np.maximum(1, np.ceil(np.expm1(self.forward(X).astype(np.float64)))).astype(int)


def set_rank_model(model):
    """Use `model` (a RankMLP, or None) for the ACA of solvers created afterwards."""
    _check(lib().rcs_set_rank_model(model.h if model is not None else None))


def set_lane(lane):
    """Library context of the calling thread (rcs_set_lane): objects created afterwards in this
    thread run on that lane's stream, so independent problems in different lanes and threads
    run concurrently on the GPU."""
    _check(lib().rcs_set_lane(int(lane)))


def set_aca_burst(steps):
    _check(lib().rcs_set_aca_burst(int(steps)))


class Operator:
    """FP64 compressed operator (assembly only) for residuals: apply(x) = Z x."""

    def __init__(self, mesh_: Mesh, bbox, leaf_size=64, tol_pow=10):
        bbox = np.ascontiguousarray(bbox, dtype=np.float64)
        h = C.c_void_p()
        _check(lib().rcs_operator_create(mesh_.h, _dp(bbox), leaf_size, tol_pow, C.byref(h)))
        self.h = h
        self.mesh = mesh_
        self.N = mesh_.N

    def __del__(self):
        if getattr(self, "h", None) and _LIB is not None:
            _LIB.rcs_operator_destroy(self.h)
            self.h = None

    def apply(self, x):
        x = np.asarray(x, dtype=np.complex128)
        single = x.ndim == 1
        Xr = x.reshape(self.N, -1)
        nrhs = Xr.shape[1]
        Xf = Xr.reshape(-1, order="F").copy()
        Y = np.zeros_like(Xf)
        _check(lib().rcs_operator_apply(self.h, _cplx_ptr(Xf), nrhs, _cplx_ptr(Y)))
        Y = Y.reshape((self.N, nrhs), order="F")
        return Y[:, 0] if single else Y


def solver_block(solver, level, box, which, key=0, cap=4_000_000):
    """One stored block of the GPU tree (which: 0 L2P, 1 P2M, 2 M2L, 3 P2P, 4 M2P, 5 P2L, 6 L2M_f)."""
    L = lib()
    L.rcs_solver_block.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int),
                                   C.POINTER(C.c_double), C.c_longlong]
    shape = np.zeros(2, dtype=np.int32)
    out = np.zeros(cap, dtype=np.complex128)
    rc = L.rcs_solver_block(solver.h, level, box, which, key, _ip(shape), _cplx_ptr(out), cap)
    if rc == 1:
        return None
    _check(rc)
    r, c = int(shape[0]), int(shape[1])
    return out[: r * c].reshape((r, c), order="F").copy()
