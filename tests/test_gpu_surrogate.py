"""GPU tests of the learned indicator surrogate: the bf16 matrix-core forward pass and
the hand-written backward pass against the CPU restatement (oracle/surrogate_ref.py),
training on estimator indicators, and the forward driver's surrogate cycles."""
import numpy as np
import pytest

from emfem_amd import api
from emfem_amd import meshgen as mg
from oracle import surrogate_ref as S

pytestmark = pytest.mark.gpu


def _spearman(a, b):
    ra = np.argsort(np.argsort(a)).astype(float)
    rb = np.argsort(np.argsort(b)).astype(float)
    ra -= ra.mean()
    rb -= rb.mean()
    return float(ra @ rb / np.sqrt((ra @ ra) * (rb @ rb)))


def test_network_matches_oracle(oracle_mod):
    m = mg.tiny_mt(n=4, aniso_form=3, seed=2)
    nb = oracle_mod.Mesh(m.points, m.tets).neighbor.astype(np.int32)
    n = nb.shape[0]
    rng = np.random.default_rng(5)
    F = rng.standard_normal((n, 16)).astype(np.float32)
    F[:, 11:] = 0.0
    target = rng.standard_normal(n).astype(np.float32)
    ctx = api.Context()
    ctx.surrogate_init(7)
    p = ctx.surrogate_params()[:S.layout()["total"]]
    y, loss, g = ctx.surrogate_eval(F, nb, target)
    yr, lr, gr = S.loss_grad(p, F.astype(np.float64), nb, target.astype(np.float64))
    assert np.allclose(y, yr, rtol=1e-4, atol=1e-4 * np.abs(yr).max()), np.abs(y - yr).max()
    assert abs(loss - lr) <= 1e-4 * lr
    L = S.layout()
    for k, size in (("w0", 16 * 64), ("b0", 64), ("w1", 128 * 64), ("b1", 64), ("w2", 128 * 64), ("w3", 128 * 64),
                    ("b3", 64), ("wo", 64)):
        a, b = g[L[k]:L[k] + size], gr[L[k]:L[k] + size]
        assert np.linalg.norm(a - b) <= 1e-2 * np.linalg.norm(b), (k, np.linalg.norm(a - b) / np.linalg.norm(b))


def _cycle_data(m, fidx=0):
    """Solve XY, YX and the dual on m's mesh; return raw features, neighbours and the
    dual's squared indicator (the factor of exy = sqrt(eta2_xy eta2_dual),
    src/em_fwd.cc:1172-1177, that the surrogate stands in for)."""
    c = api.Context("-inner_precision fp32")
    c.load_model(m)
    c.create_linear_system(3)
    c.assemble_matrix(fidx)
    c.assemble_rhs_mt(fidx, api.XY_POLAR, 0)
    c.assemble_rhs_mt(fidx, api.YX_POLAR, 1)
    c.assemble_rhs_dual(2)
    c.create_pc()
    c.solve_linear_system(0, 3, 300, 1e-8, rtols=[1e-8, 1e-8, 1e-6])
    e = c.estimate_error(fidx, [0, 1, 2])
    feats = c.surrogate_features(fidx, 0, 1)
    nb = c.topology()[2]
    return c, feats, nb, e[2]


def test_training_fits_dual_indicator():
    m = mg.adaptive_mt(cells=(8, 8, 8), n_freqs=8, n_rx_side=4)
    m.freqs = m.freqs[[3]]
    c, feats, nb, ind = _cycle_data(m)
    c.surrogate_init(1)
    loss = c.surrogate_train(feats, nb, ind, epochs=300, lr=3e-3)
    assert loss[-1] < 0.25 * loss[0], loss[[0, -1]]
    pred = c.surrogate_predict(0, 0, 1)
    assert _spearman(pred, ind) > 0.8


def test_forward_with_surrogate_cycles():
    m = mg.adaptive_mt(cells=(8, 8, 8), n_freqs=8, n_rx_side=4)
    m.freqs = m.freqs[[3]]
    m.obs = mg.mt_obs(1, len(m.rx), types=(mg.Z_XY_RI, mg.Z_YX_RI))
    c0, feats, nb, ind = _cycle_data(m)
    c0.surrogate_init(3)
    c0.surrogate_train(feats, nb, ind, epochs=300, lr=3e-3)
    params = c0.surrogate_params()
    ctx = api.Context("-max_adaptive_refinements 2 -refine_strategy fixed_number -refine_fraction 0.2 "
                      "-surrogate_cycles true -save_mesh false")
    ctx.load_model(m)
    ctx.set_observations(m.obs)
    ctx.surrogate_init(0)
    ctx.surrogate_params(params)
    ctx.forward()
    assert ctx.stat("surrogate_cycles_used") == 2
    corr = ctx.stat("surrogate_rank_corr")
    assert -1.0 <= corr <= 1.0
    assert ctx.stat("surrogate_rejected") == (1.0 if corr < 0.9 else 0.0)
    assert np.all(np.isfinite(ctx.responses()))


def _top(v, frac=0.2):
    k = max(1, int(frac * len(v)))
    return set(np.argsort(-v, kind="stable")[:k].tolist())


def test_network_gains_over_primal_only_marking():
    """What the trained network adds: on a held-out model, the elements marked with the
    primal indicator times the predicted dual factor agree better with the elements the
    full goal-oriented indicator sqrt(eta2_primal eta2_dual) marks (the reference's
    product, src/em_fwd.cc:1172-1177) than the elements the primal indicator alone marks,
    which is what marking without the dual solve and without the network would give."""
    feats, nbs, inds, off = [], [], [], 0
    for seed, f in ((2, 3), (5, 1)):
        m = mg.adaptive_mt(cells=(8, 8, 8), seed=seed, n_freqs=8, n_rx_side=4)
        m.freqs = m.freqs[[f]]
        _, fe, nb, ind = _cycle_data(m)
        feats.append(fe)
        nbs.append(np.where(nb >= 0, nb + off, -1))
        inds.append(ind)
        off += fe.shape[0]
    trainer = api.Context()
    trainer.surrogate_init(7)
    trainer.surrogate_train(np.concatenate(feats), np.concatenate(nbs), np.concatenate(inds), epochs=400, lr=3e-3)
    params = trainer.surrogate_params()
    m = mg.adaptive_mt(cells=(8, 8, 8), seed=7, n_freqs=8, n_rx_side=4)
    m.freqs = m.freqs[[2]]
    c, _, _, _ = _cycle_data(m)
    e = c.estimate_error(0, [0, 1, 2])
    c.surrogate_init(0)
    c.surrogate_params(params)
    pred = c.surrogate_predict(0, 0, 1)
    gain = []
    for pol in (0, 1):
        true = _top(np.sqrt(e[pol] * e[2]))
        with_net = len(true & _top(np.sqrt(e[pol] * pred))) / len(true)
        primal_only = len(true & _top(np.sqrt(e[pol]))) / len(true)
        print("polarisation", pol, "overlap with the goal-oriented marking: network", with_net,
              "primal only", primal_only)
        gain.append(with_net - primal_only)
    assert min(gain) > 0.05, gain
