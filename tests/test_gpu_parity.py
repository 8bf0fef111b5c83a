"""GPU parity: the HIP library (through its C ABI) against the FP64 oracle and numpy.

Bars: complex128 paths agree with the oracle to rounding (relative 1e-10 or tighter on
entries, skeletons identical); complex64 paths to FP32 rounding (relative 1e-4).
"""
import numpy as np
import pytest

import rcs_amd as R
from oracle import pyoracle as O
from rcs_amd import mesh as M

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("opA,opB", [(0, 0), (1, 0), (0, 1), (1, 1)])
@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (17, 33, 5), (64, 64, 64), (130, 70, 200),
                                   (300, 1, 700), (65, 8, 129), (100, 100, 900), (700, 650, 301)])
def test_gemm(rng, opA, opB, m, n, k):
    """Covers the 64 x 64 tile kernel (few and many tiles), the split inner index (few tiles,
    long k) and the narrow kernel (n <= 8, no transposes)."""
    A = rng.normal(size=(m, k) if opA == 0 else (k, m)) + 1j * rng.normal(size=(m, k) if opA == 0 else (k, m))
    B = rng.normal(size=(k, n) if opB == 0 else (n, k)) + 1j * rng.normal(size=(k, n) if opB == 0 else (n, k))
    ref = (A if opA == 0 else A.conj().T) @ (B if opB == 0 else B.conj().T)
    assert _rel(R.gemm(A, B, opA, opB), ref) < 1e-13
    assert _rel(R.gemm(A, B, opA, opB, precision=R.C64), ref) < 1e-5


@pytest.mark.parametrize("n,nrhs", [(1, 1), (7, 3), (33, 1), (100, 40), (549, 5), (700, 12)])
def test_lu_solve(rng, n, nrhs):
    A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)) + 2 * np.sqrt(n) * np.eye(n)
    B = rng.normal(size=(n, nrhs)) + 1j * rng.normal(size=(n, nrhs))
    X = R.lu_solve(A, B)
    assert _rel(X, np.linalg.solve(A, B)) < 1e-12
    X32 = R.lu_solve(A, B, precision=R.C64)
    assert _rel(X32, np.linalg.solve(A, B)) < 1e-4


def test_lu_partial_pivoting_needs_row_swaps(rng):
    n = 96
    A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    A[np.arange(n), np.arange(n)] = 1e-14  # tiny diagonal: every step must pivot
    B = rng.normal(size=(n, 2)) + 0j
    assert _rel(R.lu_solve(A, B), np.linalg.solve(A, B)) < 1e-11


@pytest.fixture(scope="module")
def sphere3():
    nodes, tris = M.icosphere(3, 0.15)
    return nodes, tris, O.Mesh(nodes, tris)


def test_mesh_tables_match_oracle(fixture78, sphere3):
    for nodes, tris in (fixture78, sphere3[:2]):
        g = R.Mesh(nodes, tris, R.EFIE, 1e9)
        o = O.Mesh(nodes, tris)
        t1, l1 = g.rwg()
        t2, l2 = o.rwg()
        assert np.array_equal(t1, t2) and np.array_equal(l1, l2)
        assert np.allclose(g.locations(), o.rwg_midpoints(), rtol=0, atol=1e-15)


@pytest.mark.parametrize("freq", [1e7, 1e9, 3.82e9])
def test_efie_entries_fixture78(fixture78, freq):
    nodes, tris = fixture78
    g = R.Mesh(nodes, tris, R.EFIE, freq)
    o = O.Mesh(nodes, tris)
    idx = np.arange(g.N)
    Zo = o.entries(O.EFIE, freq, idx, idx)
    Zg = g.entries(idx, idx)
    err = np.abs(Zg - Zo) / np.abs(Zo).max()
    assert err.max() < 1e-12
    Z32 = g.entries(idx, idx, precision=R.C64)
    assert _rel(Z32, Zo) < 1e-4


def test_efie_entries_sphere_blocks(sphere3, rng):
    nodes, tris, o = sphere3
    g = R.Mesh(nodes, tris, R.EFIE, 1e9)
    for _ in range(3):
        rows = rng.choice(g.N, 150, replace=False)
        cols = rng.choice(g.N, 170, replace=False)
        Zo = o.entries(O.EFIE, 1e9, rows, cols)
        assert _rel(g.entries(rows, cols), Zo) < 1e-12
    # the near-field (analytic) branch: a box of neighbours
    loc = g.locations()
    near = np.argsort(np.linalg.norm(loc - loc[0], axis=1))[:80]
    Zo = o.entries(O.EFIE, 1e9, near, near)
    assert _rel(g.entries(near, near), Zo) < 1e-12


def test_laplace_entries(sphere3):
    nodes, tris, o = sphere3
    g = R.Mesh(nodes, tris, R.LAPLACE, 1e9)
    idx = np.arange(0, g.N, 3)
    Zo = o.entries(O.LAPLACE, 1e9, idx, idx)
    assert _rel(g.entries(idx, idx), Zo) < 1e-12


def test_rhs_and_rcs(sphere3, rng):
    nodes, tris, o = sphere3
    g = R.Mesh(nodes, tris, R.EFIE, 1e9)
    kap = 2 * np.pi * 1e9 / 3e8
    for kv in ((0, 0, 1.0), (0, 0, kap), (kap * 0.6, 0, kap * 0.8)):
        assert _rel(g.rhs((1, 0, 0), kv), o.rhs(1e9, (1, 0, 0), kv)) < 1e-13
    coef = rng.normal(size=g.N) + 1j * rng.normal(size=g.N)
    th = np.linspace(0, np.pi, 19)
    ph = np.zeros_like(th)
    assert _rel(g.rcs(coef, th, ph), o.rcs(1e9, coef, th, ph)) < 1e-12


@pytest.mark.parametrize("tol", [1e-3, 1e-6])
def test_aca_skeleton_matches_oracle(sphere3, tol):
    nodes, tris, o = sphere3
    g = R.Mesh(nodes, tris, R.EFIE, 1e9)
    loc = g.locations()
    rows = np.where(loc[:, 2] > 0.12)[0]
    cols = np.where(loc[:, 2] < -0.02)[0]
    for r_, c_ in ((rows, cols), (cols, rows)):  # both branches (N1 < N2 and N1 >= N2)
        rg, rbg, cbg = g.aca(r_, c_, tol)
        ro, rbo, cbo = o.aca(O.EFIE, 1e9, r_, c_, tol)
        assert rg == ro
        assert np.array_equal(rbg, rbo) and np.array_equal(cbg, cbo)


@pytest.mark.parametrize("partition,cap", [(0, 1000), (10, 0), (10, 5), (10, 1000)])
def test_rrqr_matches_oracle(rng, partition, cap):
    m, f = 80, 60
    B = rng.normal(size=(m, 10)) + 1j * rng.normal(size=(m, 10))
    F = (rng.normal(size=(m, 12)) @ rng.normal(size=(12, f))) * (1.0 + 0j)
    F += 1e-9 * (rng.normal(size=(m, f)) + 1j * rng.normal(size=(m, f)))
    A = np.hstack([B, F])
    Qo, _, _, ro, mpo = O.rrqr(A, 6, 6, partition, 0.5, cap)
    Qg, rg, mpg = R.rrqr(A, 6, 6, partition, 0.5, cap)
    assert rg == ro
    assert abs(mpg - mpo) < 1e-12 * mpo
    assert np.linalg.norm(Qg @ Qg.conj().T - Qo @ Qo.conj().T) < 1e-10  # same kept span
    assert np.allclose(Qg.conj().T @ Qg, np.eye(rg), atol=1e-12)


def test_rrqr_early_stop_quirk():
    m = 30
    q = np.linalg.qr(np.random.default_rng(5).normal(size=(m, 3)))[0].astype(complex)
    B = np.column_stack([q[:, 0] * 3, q[:, 1] * 2, q[:, 0] * 3 + 1e-9 * q[:, 2]])
    A = np.hstack([B, np.zeros((m, 2), dtype=complex)])
    for cap, expect in ((0, 1), (5, 2)):
        Qg, rg, _ = R.rrqr(A, 6, 6, 3, 0.0, cap)
        Qo, _, _, ro, _ = O.rrqr(A, 6, 6, 3, 0.0, cap)
        assert rg == ro == expect
        assert np.linalg.norm(Qg @ Qg.conj().T - Qo @ Qo.conj().T) < 1e-10


def _solve_pair(nodes, tris, kind, freq, tol_pow, precision=R.C128, nrhs=1, rng=None):
    g = R.Mesh(nodes, tris, kind, freq)
    o = O.Mesh(nodes, tris)
    loc = g.locations()
    bb = M.bounding_box(loc)
    if kind == R.EFIE:
        b = o.rhs(freq)
    else:
        b = np.ones(g.N, dtype=complex)
    if nrhs > 1:
        b = np.column_stack([b] + [rng.normal(size=g.N) + 1j * rng.normal(size=g.N) for _ in range(nrhs - 1)])
    sg = R.AIFMM(g, bb, 64, tol_pow, tol_pow, 0, precision)
    sg.factorize()
    xg, xgb = sg.solve(b, box_order=True)
    so = O.AIFMM(o, O.EFIE if kind == R.EFIE else O.LAPLACE, freq, loc, bb, 64, tol_pow, tol_pow)
    so.factorize()
    b0 = b if b.ndim == 1 else b[:, 0]
    xo, xob = so.solve(b0, box_order=True)
    return sg, so, xg, xgb, xo, xob, b


def test_aifmm_fixture78_matches_oracle(fixture78):
    nodes, tris = fixture78
    sg, so, xg, xgb, xo, xob, b = _solve_pair(nodes, tris, R.EFIE, 1e9, 6)
    assert _rel(xg, xo) < 1e-10 and _rel(xgb, xob) < 1e-10


# (kind, freq, tol_pow, rank_slack). RRQR truncates where |R_jj| first drops under
# 10^-tol_pow times the running pivot maximum; when a diagonal sits within rounding of that
# threshold, a different (equally valid) summation order moves the cut by one column. At
# 100 MHz the EFIE operator is ill-conditioned (low-frequency breakdown), so such a cut
# moves the compressed solution by about cond * 10^-tol_pow. That case allows +-rank_slack
# per box and requires the GPU solution to be as close to the dense LU solution as the
# oracle's own compressed solution is (within a factor 2).
@pytest.mark.parametrize("kind,freq,tp,slack", [(R.LAPLACE, 1e9, 5, 0), (R.LAPLACE, 1e9, 10, 0),
                                                (R.EFIE, 1e9, 5, 0), (R.EFIE, 1e8, 7, 2)])
def test_aifmm_compressed_matches_oracle(kind, freq, tp, slack):
    """Elongated ellipsoid (real skeleton compression and RRQR fill-in recompression)."""
    nodes, tris = M.ellipsoid(3, (0.3, 0.06, 0.06))
    sg, so, xg, xgb, xo, xob, b = _solve_pair(nodes, tris, kind, freq, tp)
    st, oi = sg.stats(), so.info()
    assert (st["nLevels"], st["startLevel"]) == (oi["nLevels"], oi["startLevel"])
    assert st["n_rrqr"] == oi["n_rrqr"]
    assert st["n_compress_p2p"] == oi["n_compress_p2p"] and st["n_compress_p2l"] == oi["n_compress_p2l"]
    lv = st["nLevels"]
    rg, ro = sg.ranks(lv), so.ranks(lv)
    dr = int(np.abs(rg.astype(int) - ro.astype(int)).max())
    assert dr <= slack
    err = _rel(xg, xo)
    print(f"kind={kind} f={freq:g} tol=1e-{tp}: max rank diff {dr}, rel err vs oracle {err:.3e}")
    if dr == 0:
        assert err < 1e-8
    else:
        xd = so.mesh.dense_solve(O.EFIE if kind == R.EFIE else O.LAPLACE, freq, b)
        eg, eo = _rel(xg, xd), _rel(xo, xd)
        print(f"  vs dense LU: gpu {eg:.3e} oracle {eo:.3e}")
        assert eg < 2 * eo + 1e-12


def test_rrqr_orthonormal_lead_matches_full_rrqr(monkeypatch):
    """Fill-in recompression whose leading block is the Q of an earlier RRQR runs only the
    second partition on the projected columns; the full partitioned RRQR (RCS_FULL_RRQR=1)
    must give the same ranks and the same solution to rounding."""
    nodes, tris = M.ellipsoid(3, (0.3, 0.06, 0.06))
    out = {}
    for full in (0, 1):
        if full:
            monkeypatch.setenv("RCS_FULL_RRQR", "1")
        g = R.Mesh(nodes, tris, R.EFIE, 1e9)
        s = R.AIFMM(g, M.bounding_box(g.locations()), 64, 5, 5, 0, R.C128)
        s.factorize()
        st = s.stats()
        out[full] = (s.solve(g.rhs()), s.ranks(st["nLevels"]), st["n_compress_p2p"] + st["n_compress_p2l"])
    assert out[0][2] == out[1][2] > 0
    assert np.array_equal(out[0][1], out[1][1])
    assert _rel(out[0][0], out[1][0]) < 1e-8


def test_aifmm_multi_rhs_and_c64(rng):
    nodes, tris = M.ellipsoid(3, (0.3, 0.06, 0.06))
    sg, so, xg, xgb, xo, xob, b = _solve_pair(nodes, tris, R.LAPLACE, 1e9, 8, nrhs=4, rng=rng)
    assert _rel(xg[:, 0], xo) < 1e-8
    g = sg.mesh
    # every column solves its own right-hand side: compare with single-column solves
    for c in range(1, 4):
        assert _rel(xg[:, c], sg.solve(b[:, c])) < 1e-12
    s32 = R.AIFMM(g, M.bounding_box(g.locations()), 64, 8, 8, 0, R.C64)
    s32.factorize()
    assert _rel(s32.solve(b[:, 0]), xo) < 2e-3


def test_cpp_driver_matches_oracle(fixture78, tmp_path):
    """The reference driver's flow (test/sphere.cpp) through the C++ drop-in interface
    include/rcs_amd.hpp (tests/cpp/driver_sphere.cpp), checked against the oracle."""
    import os
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = os.path.join(root, "tests", "cpp", "driver_sphere")
    assert os.path.exists(exe), "tests/cpp/driver_sphere is built by __graft_entry__.build()"
    golden = os.path.join(root, "tests", "golden")
    out = str(tmp_path / "d")
    r = subprocess.run([exe, os.path.join(golden, "nodeList78.txt"), os.path.join(golden, "patchList78.txt"), "1e9",
                        "6", "64", out], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    load = lambda f: (lambda a: a[:, 0] + 1j * a[:, 1])(np.loadtxt(f, ndmin=2))
    phi, xbox = load(out + "_phi.txt"), load(out + "_box.txt")
    nodes, tris = fixture78
    o = O.Mesh(nodes, tris)
    loc = o.rwg_midpoints()
    b = o.rhs(1e9)
    so = O.AIFMM(o, O.EFIE, 1e9, loc, M.bounding_box(loc), 64, 6, 6)
    so.factorize()
    xo, xob = so.solve(b, box_order=True)
    assert _rel(phi, xo) < 1e-10 and _rel(xbox, xob) < 1e-10
    th_rcs = np.loadtxt(out + "_rcs.txt")
    assert _rel(th_rcs[:, 1], o.rcs(1e9, xo, th_rcs[:, 0], np.zeros(181))) < 1e-8
    n = len(xo)
    rows, cols = np.arange(n // 3), np.arange(2 * n // 3, n)
    lines = [list(map(int, ln.split())) for ln in open(out + "_aca.txt") if ln.strip()]
    ro, rbo, cbo = o.aca(O.EFIE, 1e9, rows, cols, 1e-6)
    assert lines[0] == [ro, rows.size, ro, ro, cols.size]
    bases = np.array(lines[1:], dtype=np.int64).reshape(-1, 2)
    assert np.array_equal(bases[:, 0], rbo) and np.array_equal(bases[:, 1], cbo)
    B = o.entries(O.EFIE, 1e9, rows, cols)
    _, _, _, rq, mpo = O.rrqr(B, 6, 6, 3, 0.0, 1000)
    rk, mp = np.loadtxt(out + "_rrqr.txt")
    assert int(rk) == rq and abs(mp - mpo) < 1e-10 * mpo


@pytest.mark.parametrize("kind,tp,bar", [(R.EFIE, 10, 1e-8), (R.EFIE, 6, 1e-4), (R.LAPLACE, 10, 2e-6)])
def test_operator_apply_matches_dense(rng, kind, tp, bar):
    """The FP64 compressed operator (near P2P + nested L2P M2L P2M) against the dense
    matrix-vector product of the oracle; the error follows the ACA tolerance. (For the
    Laplace collocation the incoming and outgoing skeletons differ more in size, and the
    reference's rank equalisation (T:682-689) truncates the larger one, so TOL 10 gives
    about 1e-6; the oracle's own Laplace solve at TOL 10 is 3e-5 from dense.)"""
    nodes, tris = M.ellipsoid(3, (0.3, 0.06, 0.06))
    g = R.Mesh(nodes, tris, kind, 1e9)
    o = O.Mesh(nodes, tris)
    op = R.Operator(g, M.bounding_box(g.locations()), 64, tp)
    X = rng.normal(size=(g.N, 3)) + 1j * rng.normal(size=(g.N, 3))
    Y = op.apply(X)
    ok = O.EFIE if kind == R.EFIE else O.LAPLACE
    for c in range(3):
        assert _rel(Y[:, c], o.matvec(ok, 1e9, X[:, c])) < bar


@pytest.mark.parametrize("prec,tp_factor", [(R.C64, 6), (R.C64, 4), (R.C128, 4)])
def test_refinement_reaches_dense_solution(prec, tp_factor):
    """Mixed precision: a loose (complex64 or complex128) factorization preconditions GMRES
    with the FP64 operator at TOL 12; the refined solution matches the dense LU solution of
    the oracle far below the factorization's own accuracy."""
    nodes, tris = M.ellipsoid(3, (0.3, 0.06, 0.06))
    g = R.Mesh(nodes, tris, R.EFIE, 1e9)
    o = O.Mesh(nodes, tris)
    bb = M.bounding_box(g.locations())
    b = g.rhs()
    xd = o.dense_solve(O.EFIE, 1e9, b)
    s = R.AIFMM(g, bb, 64, tp_factor, tp_factor, 0, prec)
    s.factorize()
    x0 = s.solve(b)
    op = R.Operator(g, bb, 64, 12)
    x, it, rr = s.refine(op, b, tol=1e-10, maxit=60, restart=30)
    print(f"prec={prec} TOL={tp_factor}: unrefined {_rel(x0, xd):.2e}, refined {_rel(x, xd):.2e} "
          f"after {it} GMRES iterations (relres {rr:.1e})")
    assert rr <= 1e-10
    assert _rel(x, xd) < 1e-8 and _rel(x, xd) < 1e-2 * _rel(x0, xd)


def test_refined_rcs_within_005_db_of_dense():
    """The mixed-precision form's monostatic-sweep RCS (complex64 TOL 4 factorization refined
    to a relative residual of 1e-6 against an FP64 TOL 10 operator) agrees with the RCS of the
    oracle's dense LU solution within 0.05 dB at every angle (19 angles, theta in [0, 180])."""
    nodes, tris = M.icosphere(3, 0.15)
    g = R.Mesh(nodes, tris, R.EFIE, 1e9)
    o = O.Mesh(nodes, tris)
    bb = M.bounding_box(g.locations())
    b = g.rhs()
    xd = o.dense_solve(O.EFIE, 1e9, b)
    s = R.AIFMM(g, bb, 64, 4, 4, 0, R.C64)
    s.factorize()
    op = R.Operator(g, bb, 64, 10)
    x, it, rr = s.refine(op, b, tol=1e-6, maxit=60, restart=30)
    th = np.linspace(0, np.pi, 19)
    ph = np.zeros_like(th)
    db = lambda v: 10 * np.log10(v)
    diff = np.abs(db(g.rcs(x, th, ph)) - db(g.rcs(xd, th, ph)))
    print(f"{it} GMRES iterations, relres {rr:.1e}, max RCS difference {diff.max():.2e} dB")
    assert rr <= 1e-6 and diff.max() < 0.05


def test_rank_mlp_matches_cpu_oracle(rng):
    """HIP forward and backward of the rank MLP against the numpy oracle (same BF16 weights)."""
    from oracle import rank_mlp as Q
    net = R.RankMLP(seed=3)
    p = (rng.normal(scale=0.2, size=R.MLP_NP)).astype(np.float32)
    X = rng.normal(size=(300, 8)).astype(np.float32)
    t = rng.normal(size=300).astype(np.float32)
    mean, scale = X.mean(0).astype(np.float32), X.std(0).astype(np.float32)
    net.set(p, mean, scale)
    y = net.forward(X)
    yo = Q.forward(p, mean, scale, X)
    assert np.max(np.abs(y - yo)) < 1e-4 * max(1.0, np.max(np.abs(yo)))
    loss, g = net.gradient(X, t)
    lo, go = Q.loss_and_grad(p, mean, scale, X, t)
    assert abs(loss - lo) < 1e-4 * lo
    assert np.linalg.norm(g - go) < 1e-4 * np.linalg.norm(go)


def _rank_samples(level_list, kinds=(R.EFIE,), tol=6):
    Xs, ts = [], []
    for lvl, axes in level_list:
        nodes, tris = M.ellipsoid(lvl, axes)
        for kind in kinds:
            g = R.Mesh(nodes, tris, kind, 1e9)
            s = R.AIFMM(g, M.bounding_box(g.locations()), 64, tol, tol, 0, R.C128)
            log = s.rank_log()
            Xs.append(log[:, :8])
            ts.append(np.log1p(log[:, 8]))
    return np.vstack(Xs), np.concatenate(ts)


def test_rank_model_trains_and_keeps_skeletons():
    """Train the predictor on ACA ranks logged from a few assemblies, then assemble and solve
    an unseen geometry with it: the skeletons (and so the solution) are unchanged, the ACA
    needs fewer host checks, and the miss rate (boxes whose rank exceeded the prediction)
    is reported."""
    X, t = _rank_samples([(3, (0.3, 0.06, 0.06)), (3, (0.15, 0.15, 0.15)), (3, (0.2, 0.1, 0.05))],
                         kinds=(R.EFIE, R.LAPLACE))
    net = R.RankMLP(seed=11)
    loss0 = net.gradient(X, t)[0]
    loss = net.train(X, t, iters=1500, lr=3e-3)
    assert loss < 0.25 * loss0
    nodes, tris = M.perturbed_ellipsoid(3, (0.25, 0.08, 0.08), 0.02, seed=5)
    g = R.Mesh(nodes, tris, R.EFIE, 1e9)
    bb = M.bounding_box(g.locations())
    b = g.rhs()
    s0 = R.AIFMM(g, bb, 64, 6, 6, 0, R.C128)
    st0 = s0.aca_stats()
    s0.factorize()
    x0 = s0.solve(b)
    R.set_rank_model(net)
    try:
        s1 = R.AIFMM(g, bb, 64, 6, 6, 0, R.C128)
    finally:
        R.set_rank_model(None)
    st1 = s1.aca_stats()
    s1.factorize()
    x1 = s1.solve(b)
    log1 = s1.rank_log()
    assert np.array_equal(s0.rank_log()[:, 8], log1[:, 8])  # same ACA ranks
    assert _rel(x1, x0) < 1e-13
    miss = float(np.mean(log1[:, 8] > log1[:, 9]))
    rel = np.abs(log1[:, 9] - log1[:, 8]) / np.maximum(log1[:, 8], 1)
    print(f"samples {len(X)}, loss {loss0:.3f} -> {loss:.4f}; unseen geometry: median |pred-rank|/rank "
          f"{np.median(rel):.2f}, miss rate {miss:.2f}; ACA host checks {st0['syncs']} -> {st1['syncs']}")
    assert st1["syncs"] < st0["syncs"]
    assert np.median(rel) < 0.5


@pytest.mark.skip(reason="concurrent lanes are switched off: a 64-scatterer run in 8 lanes went quiet past its "
                         "time limit and left the machine unresponsive; rcs_set_lane accepts lane 0 only")
def test_lanes_solve_independent_problems_concurrently():
    """Two problems solved at the same time from two host threads in two library lanes give
    the same solutions as one after another in the default lane."""
    import threading
    probs = [(M.ellipsoid(3, (0.3, 0.06, 0.06)), 1e9), (M.icosphere(3, 0.15), 2e9)]

    def solve(mesh, freq):
        g = R.Mesh(mesh[0], mesh[1], R.EFIE, freq)
        bb = M.bounding_box(g.locations())
        s = R.AIFMM(g, bb, 64, 5, 5, 0, R.C128)
        s.factorize()
        op = R.Operator(g, bb, 64, 8)
        x, it, rr = s.refine(op, g.rhs(), tol=1e-8, maxit=40, restart=30)
        return s.solve(g.rhs()), x

    ref = [solve(*p) for p in probs]
    out = [None, None]
    err = []

    def worker(i):
        try:
            R.set_lane(i + 1)
            out[i] = solve(*probs[i])
        except Exception as e:
            err.append(repr(e))

    ths = [threading.Thread(target=worker, args=(i,)) for i in range(2)]
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    assert not err, err
    for (x0, xr0), (x1, xr1) in zip(ref, out):
        assert _rel(x1, x0) < 1e-12 and _rel(xr1, xr0) < 1e-10


@pytest.mark.parametrize("kind", [R.LAPLACE, R.EFIE])
def test_tree_statistics_match_oracle(kind):
    """The statistics the reference driver prints through AIFMM::A, after assembly and after
    elimination, on the elongated ellipsoid at 1 GHz TOL 5 (identical ranks on both sides,
    so every stored block has the same shape)."""
    nodes, tris = M.ellipsoid(3, (0.3, 0.06, 0.06))
    g = R.Mesh(nodes, tris, kind, 1e9)
    o = O.Mesh(nodes, tris)
    loc = g.locations()
    bb = M.bounding_box(loc)
    sg = R.AIFMM(g, bb, 64, 5, 5)
    so = O.AIFMM(o, O.EFIE if kind == R.EFIE else O.LAPLACE, 1e9, loc, bb, 64, 5, 5)
    def same(stage):
        a, b = sg.tree_summary(), so.tree_summary()
        assert (a["avg_leaf_size"], a["avg_multipoles"]) == (b["avg_leaf_size"], b["avg_multipoles"]), stage
        assert abs(a["memory_gb"] - b["memory_gb"]) <= 1e-12 * b["memory_gb"], stage
        assert abs(a["memory_after_factorize_gb"] - b["memory_after_factorize_gb"]) <= \
            1e-12 * b["memory_after_factorize_gb"], stage

    same("assembled")
    sg.factorize()
    so.factorize()
    same("factorized")


def test_full_size_sphere_properties():
    """The benchmark's 30,720-unknown EFIE sphere (icosphere level 5, r = 0.15 m, 1 GHz,
    TOL 6), where the oracle is too slow to compare against, through size-independent
    properties of the solve:
    * linearity, exactly: the solution for 2b is 2x bit for bit (a power of two scales every
      product, FMA and division exactly and changes no pivot choice; the solve's reductions
      run in a fixed order);
    * right-hand sides are independent: a two-column solve gives the two single solutions;
    * the direct solution's relative residual against an FP64 TOL 10 operator sits at the
      compression level (below 1e-2), and flexible GMRES preconditioned by the factorization
      reaches 1e-6 within 10 iterations."""
    nodes, tris = M.icosphere(5, 0.15)
    g = R.Mesh(nodes, tris, R.EFIE, 1e9)
    assert g.N == 30720
    bb = M.bounding_box(g.locations())
    b = g.rhs()
    s = R.AIFMM(g, bb, 64, 6, 6)
    s.factorize()
    x = s.solve(b)
    assert np.array_equal(s.solve(2 * b), 2 * x)
    kap = 2 * np.pi * 1e9 / 3e8
    b2 = g.rhs((0.0, 1.0, 0.0), (kap, 0.0, 0.0))
    X = s.solve(np.column_stack([b, b2]))
    assert _rel(X[:, 0], x) < 1e-13 and _rel(X[:, 1], s.solve(b2)) < 1e-13
    op = R.Operator(g, bb, 64, 10)
    rr0 = _rel(op.apply(x), b)
    xr, it, rr = s.refine(op, b, tol=1e-6, maxit=30, restart=30)
    print(f"direct relres {rr0:.2e}; refined {rr:.1e} after {it} iterations")
    assert rr0 < 1e-2 and rr <= 1e-6 and it <= 10


def test_aifmm_variant_t3_matches_oracle():
    """Variant 1 (T3 parameters: eta = sqrt(3), three levels, no fill-in rank cap) on the
    elongated ellipsoid, Laplace at TOL 10, against the oracle: the same tree, and either the
    same ranks and the same solution to 1e-8, or (a cut moved by rounding at a near-tie) a
    solution as close to the dense LU solution as the oracle's."""
    nodes, tris = M.ellipsoid(3, (0.3, 0.06, 0.06))
    g = R.Mesh(nodes, tris, R.LAPLACE, 1e9)
    o = O.Mesh(nodes, tris)
    loc = g.locations()
    bb = M.bounding_box(loc)
    b = np.ones(g.N, dtype=complex)
    sg = R.AIFMM(g, bb, 64, 10, 10, 0, R.C128, 1)
    sg.factorize()
    xg = sg.solve(b)
    so = O.AIFMM(o, O.LAPLACE, 1e9, loc, bb, 64, 10, 10, 0, 1)
    so.factorize()
    xo = so.solve(b)
    st, oi = sg.stats(), so.info()
    assert (st["nLevels"], st["startLevel"]) == (oi["nLevels"], oi["startLevel"]) == (3, 2)
    dr = int(np.abs(sg.ranks(3).astype(int) - so.ranks(3).astype(int)).max())
    if dr == 0:
        assert _rel(xg, xo) < 1e-8
    else:
        xd = o.dense_solve(O.LAPLACE, 1e9, b)
        assert _rel(xg, xd) < 2 * _rel(xo, xd) + 1e-12
