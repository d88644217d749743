"""CPU checks of the surrogate's CPU restatement (oracle/surrogate_ref.py): the
analytic backward pass against central finite differences, the neighbour-mean
transpose against its definition, and bf16 rounding against known values."""
import numpy as np

from oracle import surrogate_ref as S


def _graph(n, seed):
    rng = np.random.default_rng(seed)
    nb = -np.ones((n, 4), dtype=np.int32)
    deg = np.zeros(n, dtype=int)
    for _ in range(2 * n):  # random symmetric graph of degree <= 4
        a, b = rng.integers(0, n, 2)
        if a == b or deg[a] == 4 or deg[b] == 4 or b in nb[a]:
            continue
        nb[a, deg[a]] = b; deg[a] += 1
        nb[b, deg[b]] = a; deg[b] += 1
    return nb


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 3.14159265], dtype=np.float32)
    r = S.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0  # halfway: ties to even
    assert r[2] == 1.0 + 2 ** -7       # above halfway: up
    assert r[3] == -2.5 and abs(r[4] - 3.140625) < 1e-12


def test_nb_mean_transpose():
    rng = np.random.default_rng(1)
    nb = _graph(40, 2)
    a, d = rng.standard_normal((40, 5)), rng.standard_normal((40, 5))
    # <mean(a), d> == <a, mean^T(d)>
    assert abs(np.sum(S.nb_mean(a, nb) * d) - np.sum(a * S.nb_mean_T(d, nb))) < 1e-12


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(3)
    n = 30
    nb = _graph(n, 4)
    F = rng.standard_normal((n, S.FEAT))
    F[:, 11:] = 0.0
    L = S.layout()
    p = rng.standard_normal(L["total"]) * 0.2
    target = rng.standard_normal(n)
    _, _, g = S.loss_grad(p, F, nb, target, round_bf16=False)
    idx = np.concatenate([rng.choice(np.arange(L[k], L[k] + 40), 5, replace=False)
                          for k in ("w0", "b0", "w1", "b1", "w2", "w3", "b3")] + [np.arange(L["wo"], L["wo"] + 5),
                                                                                 [L["bo"]]])
    fd = S.fd_grad(p, F, nb, target, idx)
    assert np.allclose(g[idx], fd, rtol=1e-5, atol=1e-8), np.max(np.abs(g[idx] - fd))
