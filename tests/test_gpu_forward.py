"""GPU tests of the forward driver (em_forward): receiver responses of the adaptive MT
loop and of the batched CSEM transmitters against the oracle's exact discrete solution
on the same final mesh."""
import numpy as np
import pytest

from emfem_amd import api
from emfem_amd import meshgen as mg

pytestmark = pytest.mark.gpu


def _oracle_fields(oracle_mod, m, pts, tets, attr, freq, rhs_list):
    mm = mg.Model(pts, tets, attr, m.aniso_form, m.rho, m.freqs, m.tx, m.rx, m.obs)
    om = oracle_mod.Mesh(pts, tets)
    rp, ci, C, M, _ = oracle_mod.assemble(om, mm, freq)
    sols = [oracle_mod.solve_direct(rp, ci, C, M, *r(om, mm)) for r in rhs_list]
    return om, mm, sols


@pytest.mark.parametrize("block", ["true", "false"])
def test_forward_mt_adaptive(block, oracle_mod, tmp_path):
    """One adaptive cycle, the right-hand sides solved as one block or one at a time
    (the reference's order): the final mesh's responses match the exact solution's."""
    m = mg.tiny_mt(n=3, aniso_form=1, seed=3, freq=0.5)
    ctx = api.Context("-max_adaptive_refinements 1 -refine_fraction 0.3 -e_rtol 1e-10 -inner_precision fp32 "
                      "-forward_block_rhs %s" % block)
    ctx.load_model(m)
    ctx.set_observations(m.obs)
    ctx.forward()
    rsp = ctx.responses()
    pts, tets, attr = ctx.get_mesh()
    assert tets.shape[0] > m.n_tets  # one adaptive cycle refined the mesh
    freq = m.freqs[0]
    bg_holder = {}

    def rhs(mode):
        def f(om, mm):
            if "bg" not in bg_holder:
                bg_holder["bg"] = oracle_mod.background_model(om, mm)
            return oracle_mod.rhs_mt(om, mm, bg_holder["bg"], freq, mode)
        return f

    om, mm, sols = _oracle_fields(oracle_mod, m, pts, tets, attr, freq, [rhs(-1), rhs(-2)])
    ref = oracle_mod.responses_mt(om, mm, 0, sols[0], sols[1])
    from tests.test_gpu_parity import check_mt_responses
    check_mt_responses(m.obs, rsp, ref)
    ctx.save_rsp(tmp_path / "mt.rsp")


@pytest.mark.parametrize("block", ["true", "false"])
def test_forward_csem_batched(block, oracle_mod):
    m = mg.tiny_mt(n=3, aniso_form=0, seed=1, freq=1.25)
    m.tx = np.array([[-800.0, 100.0, 400.0, 0.0, 0.0, 1.0, 0.0],
                     [0.0, -200.0, 600.0, 90.0, 0.0, 1.0, 0.0],
                     [500.0, 300.0, 300.0, 30.0, 20.0, 2.0, 600.0]])
    obs = []
    for t in range(3):
        for r in range(len(m.rx)):
            for code in (mg.F_EX_RI, mg.F_EY_RI, mg.F_HZ_RI, mg.F_EX_AP):
                obs.append((code, 0, t, r))
    m.obs = np.array(obs, dtype=np.int32)
    ctx = api.Context("-e_rtol 1e-10 -n_tx_divisions 5 -forward_block_rhs %s" % block)
    ctx.load_model(m)
    ctx.set_observations(m.obs)
    ctx.forward()
    rsp = ctx.responses()
    pts, tets, attr = ctx.get_mesh()
    freq = m.freqs[0]
    om, mm, sols = _oracle_fields(oracle_mod, m, pts, tets, attr, freq,
                                  [lambda om, mm, t=t: oracle_mod.rhs_csem(om, freq, m.tx[t], n_div=5) for t in range(3)])
    for o, (code, f, t, r) in enumerate(m.obs):
        cell = om.find_cell(m.rx[r])
        E, H = oracle_mod.interpolate_fields(om, cell, m.rx[r], freq, sols[t])
        v = oracle_mod.csem_response(code, E, H)
        # norm-wise per field at a receiver: components near zero are not resolved
        # relative to themselves by a 1e-10 residual
        scale = np.abs(E).max() if code // 10 % 10 <= 3 else np.abs(H).max()
        if code % 10 == 2:  # log amplitude + i phase
            assert abs(np.exp(rsp[o]) - np.exp(v)) <= 1e-6 * scale, (code, t, r, rsp[o], v)
        else:
            assert abs(rsp[o] - v) <= 1e-6 * scale, (code, t, r, rsp[o], v)
