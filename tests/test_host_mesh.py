"""The host-side mesh entry points of librcs_amd.so on meshes created without a device
(rcs_mesh_create_host), against the oracle: RWG numbering, unknown locations, patches,
plane-wave right-hand sides, centroid currents and RCS. These run the product library on
the CPU; none of them touches the GPU."""
import numpy as np
import pytest

import rcs_amd as R
from oracle import pyoracle as O
from rcs_amd import mesh as M


def _rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b))


@pytest.fixture(scope="module")
def meshes(fixture78):
    return [fixture78, M.icosphere(2, 0.15), M.ellipsoid(2, (0.3, 0.06, 0.06))]


def test_numbering_locations_and_patches_match_oracle(meshes):
    for nodes, tris in meshes:
        g = R.Mesh(nodes, tris, R.EFIE, 1e9, host_only=True)
        o = O.Mesh(nodes, tris)
        t1, l1 = g.rwg()
        t2, l2 = o.rwg()
        assert np.array_equal(t1, t2) and np.array_equal(l1, l2)  # bit-exact, as on the GPU mesh
        assert np.allclose(g.locations(), o.rwg_midpoints(), rtol=0, atol=1e-15)
        p, cond = g.patches()
        po = o.patches()
        assert p.shape == po.shape == (len(tris), 5)
        assert np.allclose(p, po, rtol=1e-14, atol=1e-17)
        assert np.all(cond == 0)  # no conductor column reads as conductor 0 (E:78-79)
        lap = R.Mesh(nodes, tris, R.LAPLACE, 1e9, host_only=True)
        assert lap.N == len(tris)
        assert np.allclose(lap.locations(), o.centroids(), rtol=0, atol=1e-15)


def test_rhs_current_and_rcs_match_oracle(meshes, rng):
    freq = 1e9
    kap = 2 * np.pi * freq / 3e8
    th = np.linspace(0, np.pi, 19)
    ph = np.linspace(0, 2 * np.pi, 19)
    for nodes, tris in meshes:
        g = R.Mesh(nodes, tris, R.EFIE, freq, host_only=True)
        o = O.Mesh(nodes, tris)
        for E, kv in (((1, 0, 0), (0, 0, 1.0)), ((1, 0, 0), (0, 0, kap)), ((0.8, 0, -0.6), (kap * 0.6, 0, kap * 0.8))):
            assert _rel(g.rhs(E, kv), o.rhs(freq, E, kv)) < 1e-13
        coef = rng.normal(size=g.N) + 1j * rng.normal(size=g.N)
        J = g.j_at_centroids(coef)
        assert J.shape == (len(tris), 3)
        assert _rel(J, o.j_at_centroids(freq, coef)) < 1e-13
        s = g.rcs_from_j(J, th, ph)
        assert _rel(s, o.rcs(freq, coef, th, ph)) < 1e-12
        assert np.array_equal(g.rcs(coef, th, ph), s)  # rcs_rcs is the two steps composed


def test_current_is_linear_and_rcs_quadratic(meshes, rng):
    """Size-independent properties: J is linear in the coefficients, and the RCS scales with
    the square of a common factor and is unchanged by a common phase."""
    nodes, tris = meshes[1]
    g = R.Mesh(nodes, tris, R.EFIE, 1e9, host_only=True)
    a = rng.normal(size=g.N) + 1j * rng.normal(size=g.N)
    b = rng.normal(size=g.N) + 1j * rng.normal(size=g.N)
    z = 0.3 - 1.7j
    assert _rel(g.j_at_centroids(a + z * b), g.j_at_centroids(a) + z * g.j_at_centroids(b)) < 1e-14
    th = np.linspace(0, np.pi, 7)
    ph = np.zeros_like(th)
    s = g.rcs(a, th, ph)
    assert _rel(g.rcs(2.5 * a, th, ph), 6.25 * s) < 1e-14
    assert _rel(g.rcs(np.exp(0.9j) * a, th, ph), s) < 1e-13


def test_conductor_indices_pass_through(fixture78):
    nodes, tris = fixture78
    cond = (np.arange(len(tris)) % 3).astype(np.int32)
    g = R.Mesh(nodes, tris, R.LAPLACE, 1e9, conductor=cond, host_only=True)
    assert np.array_equal(g.patches()[1], cond)


def test_host_mesh_refuses_device_work(fixture78):
    """Entries, ACA and solvers need the device tables; a host-only mesh is rejected with a
    message before any device call."""
    nodes, tris = fixture78
    g = R.Mesh(nodes, tris, R.EFIE, 1e9, host_only=True)
    with pytest.raises(RuntimeError, match="no device tables"):
        g.entries(np.arange(3), np.arange(3))
    with pytest.raises(RuntimeError, match="no device tables"):
        g.aca(np.arange(10), np.arange(50, 60), 1e-3)
    bb = M.bounding_box(g.locations())
    with pytest.raises(RuntimeError, match="no device tables"):
        R.AIFMM(g, bb, 64, 6, 6)
