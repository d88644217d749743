"""GPU parity at the plain-MT and marine-CSEM workloads' own settings, through the forward driver
(em_forward, src/em_fwd.cc:1441-1471), against the committed golden responses of the
oracle's exact discrete solutions (tests/golden/make_golden.py):

* plain MT at full size (62,581 edges, 1 Hz, 121 receivers), fp64 and fp32 inner
  cycles, the block-diagonal and the PRESB outer preconditioner, outer residual 1e-10:
  Z and T within 1e-6 (check_mt_responses);
* marine CSEM on 16 x 16 x 14 cells (27,246 edges) with all 64 x-directed dipoles as
  one block of right-hand sides at each of 0.25 / 0.75 / 1.25 Hz (forward_csem,
  src/em_fwd.cc:1382-1439), both outer preconditioners: E and H at the 29 receivers
  within 1e-6 of the largest E (or H) value of that frequency and transmitter.
"""
import os

import numpy as np
import pytest

from emfem_amd import api
from emfem_amd import meshgen as mg
from tests.golden import make_golden as G
from tests.test_gpu_parity import check_mt_responses

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _golden(name):
    z = np.load(os.path.join(GOLDEN, name))
    return z["obs"], z["rsp"], int(z["n_edges"])


def _forward(m, opts):
    c = api.Context(opts + " -max_adaptive_refinements 0 -save_mesh false")
    c.load_model(m)
    c.set_observations(m.obs)
    c.forward()
    info = c.mesh_info()
    try:
        nc = c.stat("not_converged")
    except api.EmfemError:
        nc = 0.0
    rsp = c.responses()
    c.close()
    return rsp, info, nc


@pytest.mark.parametrize("pc", ["blockdiag", "presb"])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_plain_mt_62k_edges(precision, pc):
    obs, ref, ne = _golden("plain_mt.npz")
    m = mg.plain_mt()
    assert np.array_equal(obs, m.obs)
    rsp, info, nc = _forward(m, "-inner_precision %s -e_rtol 1e-10 -A_pc_type %s" % (precision, pc))
    assert info["n_edges"] == ne and nc == 0
    check_mt_responses(obs, rsp, {o: ref[o] for o in range(len(ref))})


@pytest.mark.parametrize("pc", ["blockdiag", "presb"])
def test_marine_csem_64tx_3freq(pc):
    obs, ref, ne = _golden("marine_csem_16x16x14.npz")
    m = G.marine_csem_model()
    assert np.array_equal(obs, m.obs) and len(m.tx) == 64 and len(m.freqs) == 3
    rsp, info, nc = _forward(m, "-inner_precision fp32 -e_rtol 1e-10 -A_pc_type %s" % pc)
    assert info["n_edges"] == ne and nc == 0
    worst = 0.0
    for f in range(3):
        for t in range(64):
            sel = (obs[:, 1] == f) & (obs[:, 2] == t)
            for group in ((111, 121, 131), (141, 151, 161)):
                g = sel & np.isin(obs[:, 0], group)
                scale = np.max(np.abs(ref[g]))
                err = np.max(np.abs(rsp[g] - ref[g])) / scale
                worst = max(worst, err)
                assert err <= 1e-6, (f, t, group, err)
    print("marine CSEM (64 tx x 3 frequencies): worst receiver field error %.2e" % worst)
