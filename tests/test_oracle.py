"""Pins of the FP64 CPU oracle (oracle/) against what the mathematics fixes.

The reference ships no golden numbers (its only fixture is the 78-patch mesh), and it
cannot be compiled here (Eigen3 is absent), so each oracle function is checked against
closed forms, brute-force quadrature, physics (sphere capacitance, Mie series) or
dense LU, chosen so that a dropped term, a wrong sign or a transposed operand fails.
"""
import numpy as np
import pytest
import scipy.linalg as sla

from oracle import pyoracle as O
from rcs_amd import mesh as M
from tests.mie import pec_backscatter

EPS0 = 8.854187817e-12
C0 = 3.0e8


def test_rwg_numbering_fixture78(fixture78):
    nodes, tris = fixture78
    m = O.Mesh(nodes, tris)
    assert m.n_patch == 78 and m.n_rwg == 117  # every edge shared by two patches
    t, ln = m.rwg()
    keys = [tuple(r[:2]) for r in t]
    assert keys == sorted(keys) and all(a < b for a, b in keys)  # std::map order (E:129-209)
    for (n1, n2, tp, tm, fp, fm) in t:
        for tri, fv in ((tp, fp), (tm, fm)):
            s = set(tris[tri])
            assert {n1, n2} <= s and fv in s and fv not in (n1, n2)
        assert tp < tm  # T+ is the first patch in file order
    full = np.vstack([np.zeros((1, 3)), nodes])
    assert np.allclose(ln, np.linalg.norm(full[t[:, 0]] - full[t[:, 1]], axis=1))


def _quad_triangle(f, a, b, c, depth=6):
    """Brute force: split into 4**depth sub-triangles, 7-point Dunavant-like rule via
    centroid sampling of a further split (converges like h^2)."""
    tris = [(a, b, c)]
    for _ in range(depth):
        nt = []
        for (p, q, r) in tris:
            pq, qr, rp = (p + q) / 2, (q + r) / 2, (r + p) / 2
            nt += [(p, pq, rp), (q, qr, pq), (r, rp, qr), (pq, qr, rp)]
        tris = nt
    P = np.array([t[0] for t in tris]); Q = np.array([t[1] for t in tris]); R = np.array([t[2] for t in tris])
    area = 0.5 * np.linalg.norm(np.cross(Q - P, R - P), axis=1)
    # 3-point edge-midpoint rule (degree 2) on every sub-triangle
    pts = [(P + Q) / 2, (Q + R) / 2, (R + P) / 2]
    return sum(area[:, None] * f(x) / 3.0 for x in pts).sum(axis=0)


@pytest.mark.parametrize("h", [0.3, 1.0, 2.5])
def test_panel_integrals_vs_quadrature(rng, h):
    for _ in range(3):
        a, b, c = rng.normal(size=(3, 3))
        n = np.cross(b - a, c - a)
        n /= np.linalg.norm(n)
        size = np.linalg.norm(b - a)
        tc = (a + b + c) / 3 + rng.normal(size=3) * 0.3 * size + n * h * size
        lap, rho = O.panel(tc, a, b, c)
        cp = tc - np.dot(tc - a, n) * n  # projection of the observation point on the plane
        f1 = lambda x: (1.0 / np.linalg.norm(x - tc, axis=1))[:, None]
        fr = lambda x: (x - cp) / np.linalg.norm(x - tc, axis=1)[:, None]
        ref1 = _quad_triangle(f1, a, b, c)[0]
        refr = _quad_triangle(fr, a, b, c)
        assert abs(lap - ref1) < 2e-5 * abs(ref1)
        assert np.linalg.norm(rho - refr) < 2e-5 * np.linalg.norm(refr)


def test_panel_integral_in_plane_point():
    # observation point in the triangle's plane, outside it: 1/R integral stays finite
    a, b, c = np.array([0., 0, 0]), np.array([1., 0, 0]), np.array([0., 1, 0])
    tc = np.array([1.5, 1.2, 0.0])
    lap, _ = O.panel(tc, a, b, c)
    ref = _quad_triangle(lambda x: (1.0 / np.linalg.norm(x - tc, axis=1))[:, None], a, b, c)[0]
    assert abs(lap - ref) < 1e-5 * ref


@pytest.mark.parametrize("d,kind,tol", [(1.5, O.EFIE_PHYSICAL_PROJECTION, 1e-5), (3.0, O.EFIE_PHYSICAL_PROJECTION, 1e-5),
                                         (6.0, O.EFIE, 1e-6), (9.0, O.EFIE, 1e-6),
                                         (15.0, O.EFIE, 1e-2), (40.0, O.EFIE, 1e-2)])
def test_efie_integration_primitives_vs_quadrature(d, kind, tol):
    """getScalarIntegrationGreensFunction (E:670-685) and getVectorIntegrationGreensFunction
    (E:386-446): the integrals of G = -exp(-ikR)/(4 pi R) and of G (r' - r_free) over a source
    patch, against brute-force subdivided quadrature, at observation points d largest edges
    above the patch: the analytic singular-extraction branch (d < 4; projection on the patch's
    own plane, since the reference's origin-plane projection is not the exact integral), the
    7-point branch (4 <= d < 10; bounded by the truncated constants of E:284-286) and the
    1-point branch (d >= 10; bounded by the one-point rule itself). A wrong 7-point weight or
    point, free vertex, sign or wavenumber fails the first two bars by orders of magnitude."""
    freq = 1e9
    k = 2 * np.pi * freq / C0
    nodes, tris = M.icosphere(2, 0.15)
    o = O.Mesh(nodes, tris)
    full = np.vstack([np.zeros((1, 3)), nodes])
    P = o.patches()
    for tri in (0, 37, 101):
        v = tris[tri]
        a, b, c = full[v[0]], full[v[1]], full[v[2]]
        n = np.cross(b - a, c - a)
        n /= np.linalg.norm(n)
        tc = P[tri, :3] + n * d * P[tri, 4] + (b - a) * 0.13
        R = lambda x: np.linalg.norm(x - tc, axis=1)
        G = lambda x: (-np.exp(-1j * k * R(x)) / (4 * np.pi * R(x)))[:, None]
        ref_s = _quad_triangle(G, a, b, c, depth=7)[0]
        for fv in v:
            ref_v = _quad_triangle(lambda x: G(x) * (x - full[fv]), a, b, c, depth=7)
            s, vec = o.efie_integrals(kind, freq, tc, tri, int(fv))
            assert abs(s - ref_s) < tol * abs(ref_s), (tri, fv)
            assert np.linalg.norm(vec - ref_v) < tol * np.linalg.norm(ref_v), (tri, fv)


@pytest.mark.parametrize("lo,hi,tol", [(1.0, 4.0, 1e-5), (4.0, 10.0, 1e-6), (10.0, 30.0, 1e-2)])
def test_laplace_entries_vs_quadrature(lo, hi, tol):
    """ElectrostaticsMatrix::getMatrixEntry (src/ElectrostaticsMatrix.cpp:293-340):
    C0 * integral of 1/|c_i - r'| over patch j, C0 = 1/(4 pi eps0), against brute-force
    quadrature for patch pairs whose centroid distance lies in [lo, hi) largest edges of
    patch j: the analytic branch, the 7-point branch and the one-point branch."""
    nodes, tris = M.icosphere(3, 0.15)
    o = O.Mesh(nodes, tris)
    full = np.vstack([np.zeros((1, 3)), nodes])
    P = o.patches()
    C0 = 1.0 / (4 * np.pi * EPS0)
    cen, le = P[:, :3], P[:, 4]
    D = np.linalg.norm(cen[:, None, :] - cen[None, :, :], axis=2) / le[None, :]
    pairs = np.argwhere((D >= lo) & (D < hi))[::97][:8]
    assert len(pairs) >= 4
    Z = o.entries(O.LAPLACE, 1e9, pairs[:, 0], pairs[:, 1])
    for q, (i, j) in enumerate(pairs):
        v = tris[j]
        ref = C0 * _quad_triangle(lambda x: (1.0 / np.linalg.norm(x - cen[i], axis=1))[:, None],
                                  full[v[0]], full[v[1]], full[v[2]], depth=6)[0]
        assert abs(Z[q, q].real - ref) < tol * ref and Z[q, q].imag == 0, (i, j)


def test_laplace_capacitance_sphere():
    """Sphere of radius a: C = 4 pi eps0 a = 16.69 pF (a = 0.15 m), faceting error ~1%."""
    nodes, tris = M.icosphere(3, 0.15)
    m = O.Mesh(nodes, tris)
    N = m.n_patch
    sigma = m.dense_solve(O.LAPLACE, 1e9, np.ones(N, dtype=complex)).real
    cap = float(np.dot(sigma, m.patches()[:, 3]))  # getCapacitance, src/ElectrostaticsMatrix.cpp:366-372
    exact = 4 * np.pi * EPS0 * 0.15
    assert abs(cap - exact) / exact < 0.02


def test_mie_helper_rayleigh_limit():
    a, x = 0.15, 0.02
    assert abs(pec_backscatter(x, a) / (9 * np.pi * a * a * x ** 4) - 1) < 1e-3


@pytest.mark.parametrize("freq,kind,tol_db", [
    (3.0e8, O.EFIE, 0.2),                       # ka = 0.94: the reference's operator
    (1.0e9, O.EFIE_PHYSICAL_PROJECTION, 0.5),  # ka = 3.14 with the origin-plane quirk isolated
    (6.0e8, O.EFIE_PHYSICAL_PROJECTION, 0.5),
])
def test_efie_monostatic_rcs_vs_mie(freq, kind, tol_db):
    """Dense EFIE solve with the physical plane wave (kvec = kappa z) vs the Mie series.
    The reference projects the observation point onto the source plane shifted through
    the origin (E:469-474); that error grows with ka (8.6 dB at ka = 3.14 on this mesh),
    so the high-frequency check runs with the projection on the source plane."""
    a = 0.15
    kap = 2 * np.pi * freq / C0
    nodes, tris = M.icosphere(3, a)
    m = O.Mesh(nodes, tris)
    b = m.rhs(freq, E=(1, 0, 0), kvec=(0, 0, kap))
    x = m.dense_solve(kind, freq, b)
    s = m.rcs(freq, x, np.array([0.0]), np.array([0.0]))[0]
    mie = pec_backscatter(kap * a, a)
    assert abs(10 * np.log10(s / mie)) < tol_db


def test_aca_cur_error():
    """ACA skeleton of a well-separated block reproduces it to about the tolerance."""
    freq = 1.0e9
    nodes, tris = M.icosphere(3, 0.15)
    m = O.Mesh(nodes, tris)
    loc = m.rwg_midpoints()
    rows = np.where(loc[:, 2] > 0.135)[0]
    cols = np.where(loc[:, 2] < -0.12)[0]
    K = m.entries(O.EFIE, freq, rows, cols)
    sv = np.linalg.svd(K, compute_uv=False)
    for tol in (1e-3, 1e-5):
        r, rb, cb = m.aca(O.EFIE, freq, rows, cols, tol)
        assert 0 < r <= min(len(rows), len(cols))
        assert (sv > 10 * tol * sv[0]).sum() <= r  # no premature stop
        Ac, Ar, D = K[:, cb], K[rb, :], K[np.ix_(rb, cb)]
        err = np.linalg.norm(K - Ac @ np.linalg.solve(D, Ar)) / np.linalg.norm(K)
        assert err < 30 * tol
        assert len(set(rb)) == r and len(set(cb)) == r


def test_aca_exact_rank():
    """Rows of a block that is exactly rank one are found after one or two pivots."""
    freq = 1.0e9
    nodes, tris = M.icosphere(1, 0.15)
    m = O.Mesh(nodes, tris)
    rows = np.arange(5)
    cols = np.array([7, 7, 7, 7, 7, 7, 7, 7])  # identical columns: rank 1
    r, rb, cb = m.aca(O.EFIE, freq, rows, cols, 1e-8)
    assert r <= 2


def test_aca_stop_rule_diagonal_known_answer():
    """The stopping rule of src/ACA.cpp:702-756 worked by hand on a diagonal matrix.

    Pivots go down the diagonal; every new pair is u = d_k e_k, v = e_k, orthogonal to the
    earlier ones. The running |S| starts at 0 after the first pair (so d_0 never enters it) and
    each update adds |u|^2|v|^2 plus twice the sum over all pairs including the new one, i.e.
    3 d_k^2. With d = (1000, 1, 0.5, 0.4, 0.3) and tol = 0.3: after pair 1, 1 > 0.3*sqrt(3);
    after pair 2, 0.5 <= 0.3*sqrt(3*1.25) = 0.581, so the rank is 3. Counting d_0 would stop
    at rank 2; leaving out the newest pair's own term would run on to rank 5."""
    d = np.array([1000.0, 1.0, 0.5, 0.4, 0.3])
    K = np.zeros((5, 6), dtype=np.complex128)
    K[np.arange(5), np.arange(5)] = d
    for A, transposed in ((K, False), (K.T, True)):  # row-first (N1 < N2) and column-first branches
        r, rb, cb = O.aca_dense(A, 0.3)
        assert r == 3, (transposed, r)
        assert list(rb) == [0, 1, 2] and list(cb) == [0, 1, 2]
    # just above and below the hand-computed threshold 0.5 / sqrt(3.75) for the third pair
    thr = 0.5 / np.sqrt(3.75)
    assert O.aca_dense(K, thr * 1.001)[0] == 3
    assert O.aca_dense(K, thr * 0.999)[0] == 4
    assert O.aca_dense(K, 0.0)[0] == 5  # tol 0 runs to min(N1, N2)


def test_aca_stop_rule_cross_term_known_answer():
    """The cross terms 2|<u_l,u>||<v_l,v>| of the |S| update (src/ACA.cpp:751-753), by hand.

    K = [[10,5,0,0],[2,4,0,0],[1,0.5,0,7]]: pivot (0,0) gives v_0 = (1,.5,0,0), u_0 = (10,2,1);
    the next row is 1, its residual (0,3,0,0) pivots at column 1, so v_1 = e_1 and the residual
    column is u_1 = (0,3,0). <u_0,u_1> = 6, <v_0,v_1> = 1/2, hence
    |S|^2 = 9 + 2(6*1/2 + 9*1) = 33 and the ACA stops at rank 2 exactly when tol >= 3/sqrt(33).
    Without the cross term the threshold would be 3/sqrt(27). A global phase changes nothing."""
    K = np.array([[10, 5, 0, 0], [2, 4, 0, 0], [1, 0.5, 0, 7]], dtype=np.complex128)
    thr = 3 / np.sqrt(33)
    for A in (K, K * np.exp(0.7j)):
        for B, transposed in ((A, False), (A.T, True)):
            r, rb, cb = O.aca_dense(B, 0.55)  # between 3/sqrt(33) = 0.522 and 3/sqrt(27) = 0.577
            assert r == 2, (transposed, r)
            piv = (list(cb), list(rb)) if transposed else (list(rb), list(cb))
            assert piv == ([0, 1], [0, 1])
            assert O.aca_dense(B, thr * 1.001)[0] == 2
            assert O.aca_dense(B, thr * 0.999)[0] == 3


@pytest.mark.parametrize("m,n,r", [(30, 40, 5), (40, 30, 5), (50, 50, 2), (20, 60, 8)])
def test_aca_exact_rank_synthetic(rng, m, n, r):
    """Exact-rank recovery on a random complex rank-r matrix (r >= 2): the first r pivots
    form a nonsingular r x r intersection, and any such CUR reproduces a rank-r matrix
    exactly. The reference's rule then takes one more pair made of rounding noise (its
    |u||v| ~ 1e-16 |S| fails the test only after it is added, src/ACA.cpp:702,747), so the
    returned rank is r + 1."""
    for _ in range(3):
        A = (rng.normal(size=(m, r)) + 1j * rng.normal(size=(m, r))) @ \
            (rng.normal(size=(r, n)) + 1j * rng.normal(size=(r, n)))
        k, rb, cb = O.aca_dense(A, 1e-8)
        assert k == r + 1
        C, D, R_ = A[:, cb[:r]], A[np.ix_(rb[:r], cb[:r])], A[rb[:r], :]
        assert np.linalg.norm(A - C @ np.linalg.solve(D, R_)) < 1e-12 * np.linalg.norm(A)


def test_aca_rank_one_runs_on_noise():
    """Rank one is the degenerate case of the reference's rule: |S| never holds the first
    pair, so after it only rounding-noise pairs are compared with each other and the loop
    goes on until a noise residual vanishes or min(N1, N2) is reached. The first pivot is
    still exact."""
    rng = np.random.default_rng(11)
    a = rng.normal(size=50) + 1j * rng.normal(size=50)
    b = rng.normal(size=50) + 1j * rng.normal(size=50)
    A = np.outer(a, b)
    k, rb, cb = O.aca_dense(A, 1e-8)
    assert k >= 2
    i, j = rb[0], cb[0]
    assert np.linalg.norm(A - np.outer(A[:, j], A[i, :]) / A[i, j]) < 1e-13 * np.linalg.norm(A)


def test_aca_zero_rows_and_zero_matrix():
    """src/ACA.cpp:669-684: the first pivot row is the first row with an entry above 1e-16;
    an all-zero block has rank 0.

    Row 0 holds 1e-17 everywhere, so the scan passes it and pivots at (1, 3). Row 0 stays in
    the remaining set and, as u_0 = K[:, 3] = (1e-17, 2, 0), is the next pivot row; its
    residual (1e-17 except 0 at column 3) pivots at column 0 (first maximum). Row 2 = e_1 is
    last, at column 1. Every step is finite, and tol 1e-6 stops nothing before min(N1, N2)."""
    K = np.zeros((3, 5), dtype=np.complex128)
    assert O.aca_dense(K, 1e-6)[0] == 0
    K[0, :] = 1e-17
    K[1, 3] = 2.0
    K[2, 1] = 1.0
    r, rb, cb = O.aca_dense(K, 1e-6)
    assert r == 3
    assert list(rb) == [1, 0, 2] and list(cb) == [3, 0, 1]


def test_rrqr_plain_cpqr_matches_lapack(rng):
    A = rng.normal(size=(40, 25)) + 1j * rng.normal(size=(40, 25))
    Q, R, perm, r, mp = O.rrqr(A, 14, 14, 0, 0.0, 1000)
    _, _, piv = sla.qr(A, pivoting=True)
    assert r == 25
    assert list(perm[:r]) == list(piv[:r])  # same greedy max-norm pivot order
    assert np.allclose(Q.conj().T @ Q, np.eye(r), atol=1e-12)
    assert np.allclose(Q @ R, A[:, perm], atol=1e-11)
    assert abs(mp - np.linalg.norm(A, axis=0).max()) < 1e-12 * mp


def test_rrqr_partition_and_rank_cap(rng):
    m, p, f = 60, 10, 30
    B = np.linalg.qr(rng.normal(size=(m, p)) + 1j * rng.normal(size=(m, p)))[0]
    F = rng.normal(size=(m, f)) + 1j * rng.normal(size=(m, f))
    A = np.hstack([B, F])
    for cap in (0, 3, 7):
        Q, R, perm, r, mp = O.rrqr(A, 10, 10, p, 0.0, cap)
        assert r == p + cap
        assert set(perm[:p]) == set(range(p))  # old basis first
        # the kept span contains the old basis exactly
        assert np.linalg.norm(B - Q @ (Q.conj().T @ B)) < 1e-12
    # tolerance stop in the second partition: fill-in of numerical rank 4
    F2 = (rng.normal(size=(m, 4)) @ rng.normal(size=(4, f))).astype(complex)
    Q, R, perm, r, mp = O.rrqr(np.hstack([B, F2]), 10, 10, p, 0.0, 1000)
    assert r == p + 4


def test_rrqr_first_partition_early_stop_keeps_rank_minus_one():
    """src/rrqr.cpp:115-117: a pivot ratio below tol1 at step j sets the rank to j-1."""
    m = 30
    q = np.linalg.qr(np.random.default_rng(5).normal(size=(m, 3)))[0].astype(complex)
    # columns 0 and 2 have equal norms to double precision, so column 0 pivots first
    B = np.column_stack([q[:, 0] * 3, q[:, 1] * 2, q[:, 0] * 3 + 1e-9 * q[:, 2]])
    F = np.zeros((m, 2), dtype=complex)
    # step 2 stops (ratio ~3e-10) with rank1 = 1, not 2; with no fill-in budget the
    # second partition adds nothing, so the kept rank is 1
    Q, R, perm, r, mp = O.rrqr(np.hstack([B, F]), 6, 6, 3, 0.0, 0)
    assert r == 1
    # with budget, the second partition restarts at position 1 and re-accepts the
    # dropped direction (pivot ratio 1), then stops on the ~1e-9 column: rank 2
    Q, R, perm, r, mp = O.rrqr(np.hstack([B, F]), 6, 6, 3, 0.0, 5)
    assert r == 2
    assert np.linalg.norm(q[:, :2] - Q @ (Q.conj().T @ q[:, :2])) < 1e-12


def _orthogonal_columns(rng, m, s):
    """U diag(s) with a random unitary U (m x len(s)), columns then shuffled: pivoted QR takes
    the columns in descending s and |R_jj| = s_j exactly up to rounding."""
    U = np.linalg.qr(rng.normal(size=(m, len(s))) + 1j * rng.normal(size=(m, len(s))))[0]
    perm = rng.permutation(len(s))
    return (U * np.asarray(s))[:, perm], perm


def test_rrqr_pivot_thresholds_known_answer(rng):
    """src/rrqr.cpp:147-155 worked by hand on columns of norms s along orthogonal directions.

    Without a partition the kept rank is the number of leading pivots with
    s_j / max(inputPivot, s_0) >= tol2, capped by fillInRank. s = (1, .3, .1, .02, .005):
    tol2 = 1e-2 keeps 4; an input pivot of 3 makes the reference 3 and keeps 3 (ignoring it
    would keep 4); 50 keeps only s_0 (.3/50 < 1e-2); fillInRank 2 keeps 2."""
    s = [1.0, 0.3, 0.1, 0.02, 0.005]
    A, perm = _orthogonal_columns(rng, 12, s)
    order = list(np.argsort(perm)[:5])  # positions of s_0, s_1, ... in A
    for inp, cap, want in ((0.0, 1000, 4), (3.0, 1000, 3), (0.0, 2, 2), (0.5, 1000, 4), (50.0, 1000, 1)):
        Q, R, p, r, mp = O.rrqr(A, 2, 2, 0, inp, cap)
        assert r == want, (inp, cap, r)
        assert list(p[:r]) == order[:r]
        assert np.allclose(np.abs(np.diag(R[:, :r])), s[:r], rtol=1e-12)
        assert abs(mp - max(inp, 1.0)) < 1e-12 * max(inp, 1.0)  # the reported maxPivot


def test_rrqr_second_partition_measures_against_its_own_first_pivot(rng):
    """src/rrqr.cpp:111-118 and 146-151: the first partition's ratios use its first pivot, the
    second partition's use max(inputPivot, its own first pivot), not the first partition's.

    Partition columns of norms (1, .5), fill-in columns (.001, .0005), tol 1e-2 on both:
    the first partition keeps 2; the second partition's first pivot .001 is its own reference,
    so it keeps 2 more (rank 4). Measured against the first partition's 1 it would keep none.
    An input pivot of 1 does make the reference 1 and the rank 2."""
    m = 10
    U = np.linalg.qr(rng.normal(size=(m, 4)) + 1j * rng.normal(size=(m, 4)))[0]
    A = U * np.array([0.5, 1.0, 0.0005, 0.001])  # partition columns in ascending order on purpose
    Q, R, p, r, mp = O.rrqr(A, 2, 2, 2, 0.0, 1000)
    assert r == 4
    assert list(p[:4]) == [1, 0, 3, 2]
    assert abs(mp - 0.001) < 1e-15
    Q, R, p, r, mp = O.rrqr(A, 2, 2, 2, 1.0, 1000)
    assert r == 2 and list(p[:2]) == [1, 0]
    # a first-partition stop: norms (1, .004) with tol1 1e-2 stops at j = 1 and keeps j - 1 = 0
    # there; the second partition then restarts at position 0 over all columns with the
    # partition's unused columns zeroed, and the fill-in (.001, .0005) is measured on its own
    B = U * np.array([1.0, 0.004, 0.0005, 0.001])
    Q, R, p, r, mp = O.rrqr(B, 2, 2, 2, 0.0, 1000)
    assert r == 2


def _aifmm_vs_dense(kind, freq, tols):
    """Elongated ellipsoid: the cubic octree puts ~80-130 unknowns in each non-empty leaf,
    so skeletons are genuinely low rank and the RRQR fill-in path runs."""
    nodes, tris = M.ellipsoid(3, (0.3, 0.06, 0.06))
    m = O.Mesh(nodes, tris)
    loc = m.centroids() if kind == O.LAPLACE else m.rwg_midpoints()
    b = np.ones(len(loc), dtype=complex) if kind == O.LAPLACE else m.rhs(freq)
    xd = m.dense_solve(kind, freq, b)
    out = []
    for tp in tols:
        s = O.AIFMM(m, kind, freq, loc, M.bounding_box(loc), 64, tp, tp)
        s.factorize()
        x = s.solve(b)
        out.append((np.linalg.norm(x - xd) / np.linalg.norm(xd), s.info(), s.ranks(s.info()["nLevels"])))
    return out


def test_aifmm_laplace_converges_to_dense():
    (e5, i5, r5), (e10, i10, r10) = _aifmm_vs_dense(O.LAPLACE, 1e9, (5, 10))
    assert i5["n_rrqr"] > 0 and i10["n_rrqr"] > 0
    nz = r5[:, 2] > 0
    assert (r5[nz, 0] < r5[nz, 2]).all()  # every non-empty leaf is compressed
    assert e5 < 1e-3 and e10 < 3e-5 and e10 < e5 / 5


def test_aifmm_efie_with_compressions():
    (e5, i5, r5), = _aifmm_vs_dense(O.EFIE, 1e9, (5,))
    assert i5["n_rrqr"] > 0 and i5["n_compress_p2p"] + i5["n_compress_p2l"] > 0
    assert e5 < 1e-2


def test_rank_mlp_oracle_gradient_finite_differences():
    """The MLP oracle's analytic gradient against central differences of its loss (float64,
    without BF16 rounding so that the parameters can be perturbed continuously)."""
    from oracle import rank_mlp as Q
    rng = np.random.default_rng(7)
    p = rng.normal(scale=0.3, size=Q.NP)
    X = rng.normal(size=(40, 8))
    t = rng.normal(size=40)
    mean, scale = X.mean(0), X.std(0)
    loss, g = Q.loss_and_grad(p, mean, scale, X, t, bf16=False)
    for idx in rng.choice(Q.NP, 40, replace=False):
        e = np.zeros(Q.NP)
        e[idx] = 1e-6
        lp, _ = Q.loss_and_grad(p + e, mean, scale, X, t, bf16=False)
        lm, _ = Q.loss_and_grad(p - e, mean, scale, X, t, bf16=False)
        assert abs((lp - lm) / 2e-6 - g[idx]) < 1e-6 * max(1.0, abs(g[idx]))
    # GELU closed form: gelu(0) = 0, gelu(1) = Phi(1) = 0.8413447460685429, gelu'(0) = 0.5
    assert abs(Q.gelu(np.array([1.0]))[0] - 0.8413447460685429) < 1e-15
    assert Q.gelu(np.array([0.0]))[0] == 0.0 and abs(Q.gelu_d(np.array([0.0]))[0] - 0.5) < 1e-15
    # BF16 rounding: 1 + 2^-8 is a tie (rounds to even 1.0); 1 + 3*2^-9 rounds up to 1 + 2^-7
    assert Q.bf16_round(np.array([1.0 + 2 ** -8], np.float32))[0] == 1.0
    assert Q.bf16_round(np.array([1.0 + 3 * 2 ** -9], np.float32))[0] == np.float32(1.0 + 2 ** -7)


def test_reference_edge_map_lookup_cost_runs(fixture78):
    """The oracle's model of the reference's edge std::map (E:127-209) keeps exactly the RWG
    edges of the fixture, and both per-entry costs come back positive."""
    nodes, tris = fixture78
    m = O.Mesh(nodes, tris)
    walk, entry = m.map_walk_cost(1.0e9, 200, 3)
    assert walk > 0 and entry > 0


def test_tree_statistics_cross_checks():
    """getAvgLeafSize / getAvgMultipoles / findMemory / findMemoryAfterFactorize
    (T:490-551, 4511-4571) against other views of the same tree: the leaf average equals the
    integer mean of the point counts that ranks() reports; before elimination every P2M has
    NumMultipoles rows, so the multipole average equals the mean of NumMultipoles; and before
    elimination the two memory figures differ exactly by the interaction-list M2L blocks,
    which findMemory counts and the other does not, so the first is the larger."""
    nodes, tris = M.ellipsoid(3, (0.3, 0.06, 0.06))
    o = O.Mesh(nodes, tris)
    loc = o.rwg_midpoints()
    so = O.AIFMM(o, O.EFIE, 1e9, loc, M.bounding_box(loc), 64, 5, 5)
    inf = so.info()
    s0 = so.tree_summary()
    counts = so.ranks(inf["nLevels"])[:, 2]
    assert s0["avg_leaf_size"] == counts[counts > 0].sum() // (counts > 0).sum()
    nm = np.concatenate([so.ranks(j)[:, 0] for j in range(inf["startLevel"], inf["nLevels"] + 1)])
    assert s0["avg_multipoles"] == nm[nm > 0].sum() // (nm > 0).sum()
    assert s0["memory_gb"] > s0["memory_after_factorize_gb"] > 0
    so.factorize()
    s1 = so.tree_summary()
    assert s1["avg_leaf_size"] == s0["avg_leaf_size"]
    assert s1["memory_after_factorize_gb"] >= s0["memory_after_factorize_gb"]  # fill-ins only add blocks


def test_dense_threads_do_not_change_results(rng):
    """The oracle's dense LU and solve run on one thread by default (the reference driver's
    policy) and on several for the all-cores CPU baseline; their parallel loops are over
    independent columns, so the results are bit-identical. n = 800 puts the LU's trailing
    updates and the solve above the parallel threshold."""
    n = 800
    A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    B = rng.normal(size=(n, 24)) + 1j * rng.normal(size=(n, 24))
    old = O.set_dense_threads(1)
    try:
        X1 = O.lu_solve(A, B)
        O.set_dense_threads(8)
        X8 = O.lu_solve(A, B)
    finally:
        O.set_dense_threads(old)
    assert np.array_equal(X1, X8)
    assert np.linalg.norm(A @ X1 - B) < 1e-10 * np.linalg.norm(B)


def test_aifmm_variant_t3_converges_to_dense():
    """Variant 1 (the T3 parameters: eta = sqrt(3), at least three levels, no fill-in rank
    cap) on the elongated ellipsoid: three levels instead of two, RRQR recompressions run, and
    the solution approaches the dense LU solution as the tolerance tightens."""
    nodes, tris = M.ellipsoid(3, (0.3, 0.06, 0.06))
    m = O.Mesh(nodes, tris)
    loc = m.centroids()
    b = np.ones(len(loc), dtype=complex)
    xd = m.dense_solve(O.LAPLACE, 1e9, b)
    errs = []
    for tp in (5, 10):
        s = O.AIFMM(m, O.LAPLACE, 1e9, loc, M.bounding_box(loc), 64, tp, tp, 0, 1)
        s.factorize()
        inf = s.info()
        assert inf["nLevels"] == 3 and inf["n_rrqr"] > 0
        errs.append(np.linalg.norm(s.solve(b) - xd) / np.linalg.norm(xd))
    assert errs[0] < 5e-3 and errs[1] < 3e-5 and errs[1] < errs[0] / 10
