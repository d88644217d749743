"""Host-side helpers for the learned indicator surrogate: training data from plain
goal-oriented adaptive cycles, training, and the rank correlation used to judge it.

A training sample is one adaptive cycle of one model: the raw per-tet features
(emfem_surrogate_features), the face-neighbour graph, and the dual solution's squared
indicator eta2_dual, the factor of the goal-oriented indicator
exy = sqrt(eta2_xy * eta2_dual) (src/em_fwd.cc:1172-1177) that needs the dual solve.
The primal factors are computed exactly in every cycle, so the network only stands in
for the dual.
The cycle follows the reference's MT driver (src/em_fwd.cc:1320-1380): assemble, solve
XY / YX / dual, estimate, mark, refine.
"""
import numpy as np

from . import api


def cycle_sample(ctx, fidx=0):
    """Solve XY, YX and the dual on the context's current mesh and return
    (features, neighbours, eta2_dual, indicator exy + eyx)."""
    ctx.create_linear_system(3)
    ctx.assemble_matrix(fidx)
    ctx.assemble_rhs_mt(fidx, api.XY_POLAR, 0)
    ctx.assemble_rhs_mt(fidx, api.YX_POLAR, 1)
    ctx.assemble_rhs_dual(2)
    ctx.create_pc()
    ctx.solve_linear_system(0, 2, 300, 1e-8, check=False)
    ctx.solve_linear_system(2, 1, 300, 1e-6, check=False, precision="fp64")
    ctx.destroy_pc()
    e = ctx.estimate_error(fidx, [0, 1, 2])
    ind = np.sqrt(e[0] * e[2]) + np.sqrt(e[1] * e[2])
    feats = ctx.surrogate_features(fidx, 0, 1)
    nb = ctx.topology()[2]
    ctx.destroy_linear_system()
    return feats, nb, e[2], ind


def training_data(models, cycles=3, fraction=0.2, options="-inner_precision fp32"):
    """Samples of `cycles` plain adaptive cycles of every model (fixed-number marking
    of `fraction`), concatenated into one graph (neighbour indices offset)."""
    F, NB, Y = [], [], []
    off = 0
    for m in models:
        c = api.Context(options)
        c.load_model(m)
        for k in range(cycles):
            f, nb, dual, ind = cycle_sample(c)
            F.append(f)
            NB.append(np.where(nb >= 0, nb + off, -1))
            Y.append(dual)
            off += f.shape[0]
            if k + 1 < cycles:
                c.refine_mesh(api.refine_fixed_number(ind, fraction))
        c.close()
    return np.concatenate(F), np.concatenate(NB).astype(np.int32), np.concatenate(Y)


def spearman(a, b):
    """Rank correlation (ties broken by order; the library's check averages ties)."""
    ra = np.argsort(np.argsort(a)).astype(float)
    rb = np.argsort(np.argsort(b)).astype(float)
    ra -= ra.mean()
    rb -= rb.mean()
    return float(ra @ rb / np.sqrt((ra @ ra) * (rb @ rb)))
