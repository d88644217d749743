"""GPU checks of the solver's bookkeeping through the C ABI: the bounded acceptance at
the fp64 residual floor, captured cycles across changing column counts on one
preconditioner, the persistent solve workspace, and the split {re, im} (VecNest)
layout of the operator and preconditioner entry points (src/em_la.cc:65-111,
208-228)."""
import numpy as np
import pytest

from emfem_amd import api
from emfem_amd import meshgen as mg

pytestmark = pytest.mark.gpu


def _system(m, opts="", slots=3):
    c = api.Context(opts)
    c.load_model(m)
    c.create_linear_system(slots)
    c.assemble_matrix(0)
    c.assemble_rhs_mt(0, api.XY_POLAR, 0)
    c.assemble_rhs_mt(0, api.YX_POLAR, 1)
    c.assemble_rhs_dual(2)
    c.create_pc()
    return c


@pytest.fixture(scope="module")
def model():
    return mg.tiny_mt(n=3, aniso_form=3, seed=5, freq=0.7)


def test_rtol_below_fp64_floor_is_not_converged(model):
    """A tolerance the fp64 true residual cannot reach is reported as not converged,
    with the true residual, instead of being accepted at the floor."""
    c = _system(model)
    its, rr = c.solve_linear_system(0, 1, 300, 1e-17, check=False)
    assert rr[0] > 1e-17
    with pytest.raises(api.EmfemError) as e:
        c.solve_linear_system(0, 1, 300, 1e-17)
    assert e.value.code == api.ERR_NOT_CONVERGED
    assert "true relative residual" in str(e.value)
    # a reachable tolerance converges, and its residual is within the bound
    its, rr = c.solve_linear_system(0, 2, 300, 1e-9)
    assert np.all(rr <= 10 * 1e-9)
    c.close()


def test_graphs_across_column_counts(model):
    """1, 3, then 1 right-hand sides on one preconditioner with captured cycles give
    bitwise the plain launches' solutions: growing to 3 right-hand sides reallocates
    the cycle's buffers, and the cycles captured at 1 must not be replayed on the
    freed ones.  The plain launches run afterwards on the same preconditioner (same
    tuned kernels)."""
    c = _system(model, "-inner_precision fp32")
    sols = {}
    for graph in ("true", "false"):
        c.process_options("-use_graph %s" % graph)
        out = []
        for first, n in ((0, 1), (0, 3), (1, 1), (0, 1)):
            c.solve_linear_system(first, n, 200, 1e-9, rtols=np.full(n, 1e-9))
            out.append([c.get_solution(s) for s in range(first, first + n)])
        sols[graph] = out
    c.close()
    for a, b in zip(sols["true"], sols["false"]):
        for x, y in zip(a, b):
            assert np.array_equal(x, y)


def test_solve_workspace_is_kept(model):
    c = _system(model)
    c.solve_linear_system(0, 3, 200, 1e-8)
    w3 = c.stat("solve_workspace_bytes")
    c.solve_linear_system(0, 1, 200, 1e-8)
    c.solve_linear_system(0, 3, 200, 1e-8)
    assert c.stat("solve_workspace_bytes") == w3 > 0
    c.close()


@pytest.mark.parametrize("n_rhs", [1, 3])
def test_split_layout_matches_interleaved(model, n_rhs):
    c = _system(model, "-inner_precision fp64 -B_ksp_rtol 1e-6")
    ne = c.mesh_info()["n_edges"]
    rng = np.random.default_rng(2)
    x = rng.standard_normal((ne, 2 * n_rhs))
    xd = c.device_array(x)
    yd = c.device_array(np.zeros_like(x))
    xr, xi = c.create_vecs(n_rhs, api.LAYOUT_SPLIT)
    yr, yi = c.create_vecs(n_rhs, api.LAYOUT_SPLIT)
    try:
        assert np.all(c.copy_to_host(xr, np.ones(ne * n_rhs)) == 0.0)   # created zeroed
        c.copy_to_device(xr, np.ascontiguousarray(x[:, 0::2].T))
        c.copy_to_device(xi, np.ascontiguousarray(x[:, 1::2].T))
        for op in ("mult", "pc"):
            if op == "mult":
                c.matshell_mult_a(n_rhs, xd.ptr, yd.ptr)
                c.matshell_mult_a_split(n_rhs, xr, xi, yr, yi)
            else:
                c.pc_apply_b(n_rhs, xd.ptr, yd.ptr)
                c.pc_apply_b_split(n_rhs, xr, xi, yr, yi)
            y = yd.numpy()
            gr = c.copy_to_host(yr, np.empty(ne * n_rhs)).reshape(n_rhs, ne).T
            gi = c.copy_to_host(yi, np.empty(ne * n_rhs)).reshape(n_rhs, ne).T
            assert np.array_equal(gr, y[:, 0::2]), op
            assert np.array_equal(gi, y[:, 1::2]), op
        # in place on the halves
        c.matshell_mult_a_split(n_rhs, xr, xi, xr, xi)
        c.matshell_mult_a(n_rhs, xd.ptr, yd.ptr)
        assert np.array_equal(c.copy_to_host(xr, np.empty(ne * n_rhs)).reshape(n_rhs, ne).T, yd.numpy()[:, 0::2])
        a, b = c.create_vecs(n_rhs, api.LAYOUT_INTERLEAVED)
        assert b is None
        assert np.all(c.copy_to_host(a, np.ones(ne * 2 * n_rhs)) == 0.0)
        c.device_free(a)
    finally:
        for p in (xr, xi, yr, yi):
            c.device_free(p)
    c.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_presb_reaches_the_exact_solution(oracle_mod, precision):
    """The PRESB outer preconditioner (two inner solves with B = C + M on the
    right-hand sides' k columns) reaches the exact discrete solution that the
    reference's block-diagonal preconditioner reaches, in fewer outer iterations, and
    leaves the C ABI's block-diagonal preconditioner entry point as it is."""
    m = mg.marine_csem(cells=(10, 10, 10), n_tx=4)
    res = {}
    for pc in ("blockdiag", "presb"):
        c = api.Context("-inner_precision %s -A_pc_type %s" % (precision, pc))
        c.load_model(m)
        c.create_linear_system(4)
        c.assemble_matrix(1)
        for t in range(4):
            c.assemble_rhs_csem(1, t, t)
        c.create_pc()
        its, rr = c.solve_linear_system(0, 4, 300, 1e-10)
        assert np.all(rr <= 1e-10 * 1.0001), rr
        rp, ci, C, M = c.get_matrix()
        for t in range(4):
            e = oracle_mod.solve_direct(rp, ci, C, M, *c.get_rhs(t))
            assert np.linalg.norm(c.get_solution(t) - e) <= 1e3 * 1e-10 * np.linalg.norm(e), (pc, t)
        res[pc] = its.max()
        c.close()
    print(precision, "outer iterations", res)
    assert res["presb"] < res["blockdiag"]
