"""The committed golden receiver responses (tests/golden/, written by
tests/golden/make_golden.py from the oracle's exact discrete solutions) agree with a
fresh oracle computation: all of the plain-MT responses, and for the marine-CSEM
model one frequency at three transmitters."""
import os

import numpy as np

from tests.golden import make_golden as G

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    z = np.load(os.path.join(HERE, name))
    return z["obs"], z["rsp"], int(z["n_edges"])


def test_plain_mt_golden(oracle_mod):
    obs, rsp, ne = load("plain_mt.npz")
    o2, r2, ne2 = G.plain_mt_responses()
    assert ne == ne2 == 62581
    assert np.array_equal(obs, o2)
    assert np.max(np.abs(rsp - r2)) <= 1e-12 * np.max(np.abs(r2))


def test_marine_csem_golden_sample(oracle_mod):
    O = oracle_mod
    obs, rsp, ne = load("marine_csem_16x16x14.npz")
    m = G.marine_csem_model()
    om = O.Mesh(m.points, m.tets)
    assert om.n_edges == ne and np.array_equal(obs, m.obs)
    fidx = 1
    f = float(m.freqs[fidx])
    rp, ci, C, M, _ = O.assemble(om, m, f)
    solve = O.direct_solver(rp, ci, C, M)
    for t in (0, 31, 63):
        e = solve(*O.rhs_csem(om, f, m.tx[t]))
        for o in np.nonzero((obs[:, 1] == fidx) & (obs[:, 2] == t))[0]:
            code, _, _, r = obs[o]
            E, H = O.interpolate_fields(om, om.find_cell(m.rx[r]), m.rx[r], f, e)
            v = O.csem_response(int(code), E, H)
            assert abs(v - rsp[o]) <= 1e-12 * max(abs(v), 1e-300) + 1e-300, (code, t, r)
