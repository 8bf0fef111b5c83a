"""Training data and training of the learned ACA rank predictor.

Every assembly logs, per ACA problem (a box's incoming or outgoing skeleton at one level),
the predictor's 8 features and the rank the adaptive ACA found (AIFMM.rank_log()). A model
is trained on such logs (target log(1 + rank)) with the library's own HIP kernels and then
handed to the library with set_rank_model(): solvers created afterwards size their ACA
buffers and launch their first ACA steps from the predicted ranks; the adaptive stop
still decides where each ACA ends.
"""
import numpy as np

from . import AIFMM, EFIE, LAPLACE, C128, Mesh, RankMLP
from . import mesh as M


def samples(solver):
    """(X, t) from one assembled solver: features and log(1 + rank)."""
    log = solver.rank_log()
    return log[:, :8].astype(np.float32), np.log1p(log[:, 8]).astype(np.float32)


def assemble_and_log(nodes, tris, kind, freq, tol_pow, leaf_size=64):
    g = Mesh(nodes, tris, kind, freq)
    s = AIFMM(g, M.bounding_box(g.locations()), leaf_size, tol_pow, tol_pow, 0, C128)
    return samples(s)


def training_configs(seed=2026, n_ellipsoids=6):
    """Small instances of the workload families the predictor is trained on: the Laplace
    sphere (5,120 unknowns, TOL 8), the EFIE sphere at 1 GHz (7,680 unknowns at TOL 4, 6
    and 8; 30,720 unknowns at TOL 4) and at 3.82 GHz (7,680 unknowns at TOL 4 and 6, the
    electrical box sizes of the 122,880-unknown sphere's coarse levels), and seeded
    ellipsoids with semi-axes in [0.05, 0.3] m at 1 GHz (subdivision 3, TOL 6)."""
    cfg = [("sphere", 4, (0.15, 0.15, 0.15), LAPLACE, 1e9, 8)]
    cfg += [("sphere", 4, (0.15, 0.15, 0.15), EFIE, 1e9, tp) for tp in (4, 6, 8)]
    cfg += [("sphere", 5, (0.15, 0.15, 0.15), EFIE, 1e9, 4)]
    cfg += [("sphere", 4, (0.15, 0.15, 0.15), EFIE, 3.82e9, tp) for tp in (4, 6)]
    rng = np.random.default_rng(seed)
    for _ in range(n_ellipsoids):
        axes = tuple(float(a) for a in rng.uniform(0.05, 0.3, size=3))
        cfg.append(("ellipsoid", 3, axes, EFIE, 1e9, 6))
    return cfg


def collect(configs):
    Xs, ts = [], []
    for shape, level, axes, kind, freq, tol in configs:
        nodes, tris = M.ellipsoid(level, axes) if shape == "ellipsoid" else M.icosphere(level, axes[0])
        X, t = assemble_and_log(nodes, tris, kind, freq, tol)
        Xs.append(X)
        ts.append(t)
    return np.vstack(Xs), np.concatenate(ts)


def train(X, t, seed=1, iters=2000, lr=3e-3):
    net = RankMLP(seed=seed)
    loss0 = net.gradient(X, t)[0]
    loss = net.train(X, t, iters=iters, lr=lr)
    return net, loss0, loss
