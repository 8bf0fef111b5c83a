"""Synthetic surface meshes in the reference's input format.

The reference reads 1-based ``nodeList<N>.txt`` ("x y z" per line) and
``patchList<N>.txt`` ("i j k" per line) files (src/EFIEMatrix.cpp:31-95,
test/sphere.cpp:64-65). The workloads here are generated: icospheres, ellipsoids and
radially perturbed ellipsoids, written and read in that same format.
"""
from __future__ import annotations

import numpy as np

__all__ = [
    "icosphere",
    "ellipsoid",
    "perturbed_ellipsoid",
    "write_reference_files",
    "read_reference_files",
    "bounding_box",
]


def _icosahedron():
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = np.array(
        [[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0],
         [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
         [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    f = np.array(
        [[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
         [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
         [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
         [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], dtype=np.int64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    return v, f


def icosphere(level: int, radius: float = 0.15):
    """Icosahedron subdivided ``level`` times (20*4**level flat triangles), vertices
    projected to ``radius``. Returns (nodes (n,3) float64, tris (m,3) int32, 1-based)."""
    v, f = _icosahedron()
    verts = [tuple(p) for p in v]
    for _ in range(level):
        cache = {}
        new_f = []

        def mid(a, b):
            key = (a, b) if a < b else (b, a)
            if key in cache:
                return cache[key]
            p = (np.asarray(verts[a]) + np.asarray(verts[b])) / 2.0
            p = p / np.linalg.norm(p)
            verts.append(tuple(p))
            cache[key] = len(verts) - 1
            return cache[key]

        for a, b, c in f:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            new_f += [[a, ab, ca], [b, bc, ab], [c, ca, bc], [ab, bc, ca]]
        f = np.asarray(new_f, dtype=np.int64)
    nodes = np.asarray(verts, dtype=np.float64) * radius
    return nodes, (f + 1).astype(np.int32)


def ellipsoid(level: int, axes):
    """Icosphere of unit radius scaled by the semi-axes (a, b, c)."""
    nodes, tris = icosphere(level, 1.0)
    return nodes * np.asarray(axes, dtype=np.float64)[None, :], tris


def perturbed_ellipsoid(level: int, axes, noise: float, seed: int):
    """Ellipsoid with seeded radial noise of relative amplitude ``noise``."""
    nodes, tris = ellipsoid(level, axes)
    rng = np.random.default_rng(seed)
    scale = 1.0 + noise * rng.uniform(-1.0, 1.0, size=(nodes.shape[0], 1))
    return nodes * scale, tris


def write_reference_files(nodes, tris, node_path, patch_path, conductor=None):
    """Node lines "x y z"; patch lines "i j k" (1-based), or "i j k c" with a conductor index."""
    np.savetxt(node_path, nodes, fmt="%.18e")
    t = np.asarray(tris)
    if conductor is not None:
        t = np.column_stack([t, np.asarray(conductor)])
    np.savetxt(patch_path, t, fmt="%d")


def read_reference_files(node_path, patch_path, with_conductor=False):
    """Inverse of write_reference_files. A patch file without the fourth column gives
    conductor 0 for every patch, as the reference's reader does (src/EFIEMatrix.cpp:72-95)."""
    nodes = np.loadtxt(node_path, dtype=np.float64).reshape(-1, 3)
    tris = np.loadtxt(patch_path, dtype=np.int64).reshape(-1, np.loadtxt(patch_path, max_rows=1).size)
    t3 = tris[:, :3].astype(np.int32)
    if not with_conductor:
        return nodes, t3
    cond = tris[:, 3].astype(np.int32) if tris.shape[1] > 3 else np.zeros(len(tris), dtype=np.int32)
    return nodes, t3, cond


def bounding_box(points):
    """(xmin, xmax, ymin, ymax, zmin, zmax) of the unknowns' locations (S:159-174)."""
    mn, mx = points.min(axis=0), points.max(axis=0)
    return np.array([mn[0], mx[0], mn[1], mx[1], mn[2], mx[2]], dtype=np.float64)
