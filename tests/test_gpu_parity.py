"""GPU parity: the HIP path through the C ABI against the CPU FP64 oracle.

Bars: topology/pattern bit-exact; assembled C and M within 1e-12 (relative to the
largest entry); right-hand sides within 1e-12; FGMRES solutions within the
discretisation-independent bound set by the requested residual; receiver responses
within 1e-6 relative of the responses of the exact discrete solution.
"""
import numpy as np
import pytest

from emfem_amd import api
from emfem_amd import meshgen as mg

pytestmark = pytest.mark.gpu


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    scale = np.max(np.abs(b)) if b.size else 1.0
    return float(np.max(np.abs(a - b)) / scale) if b.size else 0.0


@pytest.fixture(scope="module")
def tiny(oracle_mod):
    m = mg.tiny_mt(n=3, aniso_form=3, seed=5, freq=0.7)
    om = oracle_mod.Mesh(m.points, m.tets)
    return m, om


def _ctx(m, opts=""):
    c = api.Context(opts)
    c.load_model(m)
    return c


def test_assembly_matches_oracle(tiny, oracle_mod):
    m, om = tiny
    ctx = _ctx(m)
    ctx.create_linear_system(3)
    ctx.assemble_matrix(0)
    rp, ci, C, M = ctx.get_matrix()
    orp, oci = om.pattern()
    assert np.array_equal(rp, orp) and np.array_equal(ci, oci), "CSR pattern differs"
    _, _, oC, oM, _ = oracle_mod.assemble(om, m, m.freqs[0], (orp, oci))
    assert _rel(C, oC) < 1e-12
    assert _rel(M, oM) < 1e-12
    # bitwise reproducible (colour-ordered, atomic-free scatter)
    ctx.assemble_matrix(0)
    _, _, C2, M2 = ctx.get_matrix()
    assert np.array_equal(C, C2) and np.array_equal(M, M2)


@pytest.mark.parametrize("form", [0, 1, 2, 3])
def test_assembly_all_aniso_forms(form, oracle_mod):
    m = mg.tiny_mt(n=2, aniso_form=form, seed=11, freq=3.0)
    om = oracle_mod.Mesh(m.points, m.tets)
    ctx = _ctx(m)
    ctx.create_linear_system(1)
    ctx.assemble_matrix(0)
    rp, ci, C, M = ctx.get_matrix()
    _, _, oC, oM, _ = oracle_mod.assemble(om, m, m.freqs[0], (rp, ci))
    assert _rel(C, oC) < 1e-12 and _rel(M, oM) < 1e-12


def test_operator_matches_oracle(tiny, oracle_mod):
    m, om = tiny
    ctx = _ctx(m)
    ctx.create_linear_system(1)
    ctx.assemble_matrix(0)
    rp, ci, C, M = ctx.get_matrix()
    ne = rp.size - 1
    rng = np.random.default_rng(3)
    for k in (1, 3, 8, 20, 70):
        x = rng.standard_normal((ne, 2 * k))
        xd = ctx.device_array(x)
        yd = ctx.device_array(np.zeros_like(x))
        ctx.matshell_mult_a(k, xd.ptr, yd.ptr)
        y = yd.numpy()
        for j in range(k):
            yr, yi = oracle_mod.apply_operator(rp, ci, C, M, x[:, 2 * j], x[:, 2 * j + 1])
            assert _rel(y[:, 2 * j], yr) < 1e-13
            assert _rel(y[:, 2 * j + 1], yi) < 1e-13


def test_rhs_matches_oracle(tiny, oracle_mod):
    m, om = tiny
    m = mg.tiny_mt(n=3, aniso_form=3, seed=5, freq=0.7)
    m.tx = np.array([[100.0, -200.0, 500.0, 30.0, 10.0, 2.0, 0.0],
                     [0.0, 0.0, 800.0, 0.0, 0.0, 1.0, 900.0]])
    ctx = _ctx(m, "-n_tx_divisions 7")
    ctx.create_linear_system(5)
    bg = oracle_mod.background_model(om, m)
    for slot, mode in ((0, api.XY_POLAR), (1, api.YX_POLAR)):
        ctx.assemble_rhs_mt(0, mode, slot)
        re, im = ctx.get_rhs(slot)
        ore, oim = oracle_mod.rhs_mt(om, m, bg, m.freqs[0], mode)
        assert _rel(re, ore) < 1e-12 and _rel(im, oim) < 1e-12
    ctx.assemble_rhs_dual(2)
    re, im = ctx.get_rhs(2)
    ore, oim = oracle_mod.rhs_dual(om, m.rx)
    assert _rel(re, ore) < 1e-12 and _rel(im, oim) < 1e-12
    for t in range(2):
        ctx.assemble_rhs_csem(0, t, 3 + t)
        re, im = ctx.get_rhs(3 + t)
        ore, oim = oracle_mod.rhs_csem(om, m.freqs[0], m.tx[t], n_div=7)
        assert np.all(re == 0.0)
        assert _rel(im, oim) < 1e-12


def check_mt_responses(obs, rsp, ref, tol=1e-6):
    """Per receiver: the impedance tensor within tol of its largest entry (Z_xx and
    Z_yy are differences of near-equal products), the tipper within tol absolute
    (dimensionless)."""
    for r in np.unique(obs[:, 3]):
        rows = [o for o in ref if obs[o][3] == r]
        zrows = [o for o in rows if obs[o][0] in (311, 321, 331, 341)]
        scale = max(abs(ref[o]) for o in zrows) if zrows else 1.0
        for o in rows:
            s = scale if o in zrows else 1.0
            assert abs(rsp[o] - ref[o]) <= tol * s, (obs[o], rsp[o], ref[o])


@pytest.mark.parametrize("precision,rtol,rsp_tol,cycle", [("fp64", 1e-10, 1e-6, "additive"),
                                                           ("fp32", 1e-10, 1e-6, "additive"),
                                                           ("fp64", 1e-8, 2e-5, "additive"),
                                                           ("fp32", 1e-8, 2e-5, "additive"),
                                                           ("fp32", 1e-10, 1e-6, "multiplicative"),
                                                           ("fp64", 1e-8, 2e-5, "multiplicative")])
def test_mt_solve_and_responses(precision, rtol, rsp_tol, cycle, oracle_mod):
    """The FP32 inner solve inside FP64 FGMRES reaches the same true residual as the
    plain FP64 path; responses then agree with the exact discrete solution to a bound
    set by that residual (1e-6 at 1e-10; at 1e-8 the solution error is ~1e-6..1e-5
    relative on this mesh).  Both auxiliary cycles (0-(G+Pi)-0 and 0-G-Pi-G-0)."""
    m = mg.halfspace_mt(n=4, n_rx=3)
    om = oracle_mod.Mesh(m.points, m.tets)
    ctx = _ctx(m, "-inner_precision %s -ams_cycle %s" % (precision, cycle))
    ctx.create_linear_system(3)
    ctx.assemble_matrix(0)
    ctx.assemble_rhs_mt(0, api.XY_POLAR, 0)
    ctx.assemble_rhs_mt(0, api.YX_POLAR, 1)
    ctx.assemble_rhs_dual(2)
    ctx.create_pc()
    its, rr = ctx.solve_linear_system(0, 3, 300, rtol, check=False)
    print(precision, cycle, "outer its", its, "true relres", rr, "inner its", ctx.stat("inner_its"))
    assert np.all(rr <= rtol * 1.0001), rr
    rp, ci, C, M = ctx.get_matrix()
    sols = []
    for slot in range(3):
        sr, si = ctx.get_rhs(slot)
        e_direct = oracle_mod.solve_direct(rp, ci, C, M, sr, si)
        e_gpu = ctx.get_solution(slot)
        assert np.linalg.norm(e_gpu - e_direct) <= 1e3 * rtol * np.linalg.norm(e_direct)
        sols.append(e_direct)
    rsp = ctx.calculate_response_mt(0, 0, 1, m.obs)
    ref = oracle_mod.responses_mt(om, m, 0, sols[0], sols[1])
    check_mt_responses(m.obs, rsp, ref, tol=rsp_tol)
    # apparent resistivity of a half-space approaches its resistivity (coarse mesh: within 40%)
    zxy = rsp[[o for o in range(len(m.obs)) if m.obs[o][0] == mg.Z_XY_RI]]
    rho_a = np.abs(zxy) ** 2 / (2 * np.pi * m.freqs[0] * 4e-7 * np.pi)
    assert np.all(np.abs(rho_a / 100.0 - 1) < 0.4)


def test_estimator_matches_oracle(oracle_mod):
    m = mg.tiny_mt(n=3, aniso_form=2, seed=7, freq=2.0)
    om = oracle_mod.Mesh(m.points, m.tets)
    ctx = _ctx(m, "-inner_precision fp64")
    ctx.create_linear_system(3)
    ctx.assemble_matrix(0)
    ctx.assemble_rhs_mt(0, api.XY_POLAR, 0)
    ctx.assemble_rhs_mt(0, api.YX_POLAR, 1)
    ctx.assemble_rhs_dual(2)
    ctx.create_pc()
    ctx.solve_linear_system(0, 3, 200, 1e-10)
    sols = [ctx.get_solution(s) for s in range(3)]
    err = ctx.estimate_error(0, [0, 1, 2])
    oerr = oracle_mod.estimate_error(om, m, m.freqs[0], sols)
    for s in range(3):
        assert _rel(err[s], oerr[s]) < 1e-12
    # the per-tet indicator the reference refines on (src/em_fwd.cc:1172-1177)
    ixy = np.sqrt(err[0] * err[2])
    oxy = np.sqrt(oerr[0] * oerr[2])
    assert _rel(ixy, oxy) < 1e-12


@pytest.mark.parametrize("n_rhs", [1, 2, 3])
def test_pc_apply_matches_fcg_tolerance(tiny, n_rhs, oracle_mod):
    """pc_apply_b gives x with ||b - B x|| <= rtol ||b|| per column (src/em_la.cc:208-228),
    on 2, 4 and 6 real columns (6: the padded reduce-scatter of the SpMMs), with the
    first-kind Chebyshev smoother and with the fourth-kind one."""
    m, om = tiny
    for prec, tol, kind in (("fp64", 1e-2, "first"), ("fp32", 1e-2, "first"), ("fp64", 1e-6, "first"),
                            ("fp32", 1e-2, "fourth"), ("fp64", 1e-6, "fourth")):
        ctx = _ctx(m, "-inner_precision %s -B_ksp_rtol %g -ams_cheb_kind %s" % (prec, tol, kind))
        ctx.create_linear_system(1)
        ctx.assemble_matrix(0)
        ctx.create_pc()
        rp, ci, C, M = ctx.get_matrix()
        B = oracle_mod.csr(rp, ci, C + M, rp.size - 1)
        rng = np.random.default_rng(9)
        b = rng.standard_normal((rp.size - 1, 2 * n_rhs))
        bd = ctx.device_array(b)
        xd = ctx.device_array(np.zeros_like(b))
        ctx.pc_apply_b(n_rhs, bd.ptr, xd.ptr)
        x = xd.numpy()
        # the per-column stopping test ends the FCG once the tolerance is met: far
        # fewer than B_ksp_max_it iterations, and not far below the tolerance
        assert ctx.stat("inner_its") < (30 if tol >= 1e-2 else 100), ctx.stat("inner_its")
        for c in range(2 * n_rhs):
            r = b[:, c] - B @ x[:, c]
            assert np.linalg.norm(r) <= tol * np.linalg.norm(b[:, c]) * 1.01
            assert np.linalg.norm(r) >= 1e-3 * tol * np.linalg.norm(b[:, c])


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
@pytest.mark.parametrize("cycle", ["additive", "multiplicative"])
def test_structured_transfers_match_csr(tiny, precision, cycle):
    """The structured auxiliary-space transfers (edge endpoints and half tangents) apply
    the same Pi and Pi^T as the CSR matrices of the interpolation: one FCG step
    (x = alpha M b) agrees to the cycle's rounding, on 2 and 6 real columns.  The
    gradient space is nearly singular in the air (only the tiny mass term), so the
    auxiliary solves amplify rounding differences (fp64: ~1e-10 relative here); a wrong
    entry of Pi would differ at O(1)."""
    m, om = tiny
    ctx = _ctx(m, "-inner_precision %s -B_ksp_max_it 1 -ams_cycle %s" % (precision, cycle))
    ctx.create_linear_system(1)
    ctx.assemble_matrix(0)
    ne = ctx.mesh_info()["n_edges"]
    rng = np.random.default_rng(21)
    tol = 1e-4 if precision == "fp32" else 1e-8
    for n_rhs in (1, 3):
        b = rng.standard_normal((ne, 2 * n_rhs))
        bd = ctx.device_array(b)
        out = {}
        for tr in ("csr", "structured"):
            ctx.process_options("-ams_transfer %s" % tr)
            ctx.create_pc()
            xd = ctx.device_array(np.zeros_like(b))
            ctx.pc_apply_b(n_rhs, bd.ptr, xd.ptr)
            out[tr] = xd.numpy()
        assert np.all(np.isfinite(out["structured"]))
        for c in range(2 * n_rhs):
            d = np.linalg.norm(out["structured"][:, c] - out["csr"][:, c])
            rel = d / np.linalg.norm(out["csr"][:, c])
            print("transfer", precision, cycle, n_rhs, c, rel)
            assert rel <= tol, (n_rhs, c, rel)


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_structured_transfers_match_csr_on_3_columns(tiny, precision, oracle_mod):
    """On 3 real columns (the PRESB halves of three right-hand sides) the structured
    prolongation runs a lane per edge for all columns: the solves with the structured
    transfers and with the CSR matrices of Pi both reach the exact discrete solution."""
    m, om = tiny
    for tr in ("csr", "structured"):
        ctx = _ctx(m, "-inner_precision %s -A_pc_type presb -ams_transfer %s" % (precision, tr))
        ctx.create_linear_system(3)
        ctx.assemble_matrix(0)
        ctx.assemble_rhs_mt(0, api.XY_POLAR, 0)
        ctx.assemble_rhs_mt(0, api.YX_POLAR, 1)
        ctx.assemble_rhs_mt(0, api.XY_POLAR, 2)
        ctx.create_pc()
        its, rr = ctx.solve_linear_system(0, 3, 300, 1e-10)
        assert np.all(rr <= 1e-10 * 1.0001), (tr, rr)
        rp, ci, C, M = ctx.get_matrix()
        for s in range(3):
            e = oracle_mod.solve_direct(rp, ci, C, M, *ctx.get_rhs(s))
            g = ctx.get_solution(s)
            assert np.linalg.norm(g - e) <= 1e4 * 1e-10 * np.linalg.norm(e), (tr, s)
        ctx.close()


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
@pytest.mark.parametrize("cycle", ["additive", "multiplicative"])
def test_device_setup_matches_host(tiny, precision, cycle):
    """The auxiliary-space Galerkin products and the l1 scaling formed on the device
    (the default set-up) give the preconditioner of the host set-up: one FCG step agrees
    to the cycle's rounding (the products are summed in the host's order; the kernels'
    row mappings are tuned per set-up and may differ), on 2 and 6 real columns."""
    m, om = tiny
    ctx = _ctx(m, "-inner_precision %s -B_ksp_max_it 1 -ams_cycle %s" % (precision, cycle))
    ctx.create_linear_system(1)
    ctx.assemble_matrix(0)
    ne = ctx.mesh_info()["n_edges"]
    rng = np.random.default_rng(23)
    tol = 1e-4 if precision == "fp32" else 1e-8
    for n_rhs in (1, 3):
        b = rng.standard_normal((ne, 2 * n_rhs))
        bd = ctx.device_array(b)
        out = {}
        for where in ("host", "device"):
            ctx.process_options("-ams_setup %s" % where)
            ctx.create_pc()
            xd = ctx.device_array(np.zeros_like(b))
            ctx.pc_apply_b(n_rhs, bd.ptr, xd.ptr)
            out[where] = xd.numpy()
        assert np.all(np.isfinite(out["device"]))
        for c in range(2 * n_rhs):
            rel = np.linalg.norm(out["device"][:, c] - out["host"][:, c]) / np.linalg.norm(out["host"][:, c])
            print("setup", precision, cycle, n_rhs, c, rel)
            assert rel <= tol, (n_rhs, c, rel)


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_graph_replay_matches_plain_launches(tiny, precision):
    """The captured auxiliary-space cycle gives bitwise the launches' result on every
    replay, with eager work and synchronisation between replays."""
    m, om = tiny
    ctx = _ctx(m, "-inner_precision %s -B_ksp_rtol 1e-6" % precision)
    ctx.create_linear_system(2)
    ctx.assemble_matrix(0)
    ctx.create_pc()
    rng = np.random.default_rng(12)
    ne = ctx.mesh_info()["n_edges"]
    b = rng.standard_normal((ne, 4))
    bd = ctx.device_array(b)
    outs = {}
    for graph in ("false", "true"):
        ctx.process_options("-use_graph %s" % graph)
        res = []
        for rep in range(3):
            xd = ctx.device_array(np.zeros_like(b))
            ctx.pc_apply_b(2, bd.ptr, xd.ptr)
            res.append(xd.numpy())
            ctx.matshell_mult_a(2, bd.ptr, xd.ptr)  # eager work between replays
        outs[graph] = res
    for r in outs["true"] + outs["false"][1:]:
        assert np.array_equal(r, outs["false"][0])


def test_not_converged_is_reported(oracle_mod):
    m = mg.tiny_mt(n=3)
    ctx = _ctx(m, "-B_ksp_max_it 1")
    ctx.create_linear_system(1)
    ctx.assemble_matrix(0)
    ctx.assemble_rhs_mt(0, api.XY_POLAR, 0)
    ctx.create_pc()
    with pytest.raises(api.EmfemError) as e:
        ctx.solve_linear_system(0, 1, 1, 1e-12)
    assert e.value.code == api.ERR_NOT_CONVERGED


def test_device_point_location_matches_oracle(oracle_mod):
    """emfem_locate_points (device nearest vertex + breadth-first search) returns the
    oracle's tet (src/em_mesh.cc:518-566) for receivers on the surface, vertices, edge
    midpoints, face centroids (points on shared faces, where the search order decides),
    random interior points and points outside the mesh."""
    m = mg.adaptive_mt(cells=(8, 8, 8), n_rx_side=6)
    om = oracle_mod.Mesh(m.points, m.tets)
    ctx = _ctx(m)
    rng = np.random.default_rng(21)
    pts = m.points
    tets = m.tets
    face_c = pts[tets[:, [1, 2, 3]]].mean(axis=1)[rng.choice(len(tets), 200, replace=False)]
    edge_m = 0.5 * (pts[tets[:, 0]] + pts[tets[:, 1]])[rng.choice(len(tets), 200, replace=False)]
    lo, hi = pts.min(0), pts.max(0)
    inner = lo + (hi - lo) * rng.random((300, 3))
    outside = np.array([[hi[0] + 10.0, 0.0, 0.0], [0.0, lo[1] - 1.0, 0.0], [0.0, 0.0, hi[2] + 5.0]])
    q = np.concatenate([m.rx, pts[rng.choice(len(pts), 200, replace=False)], edge_m, face_c, inner, outside])
    got = ctx.locate_points(q)
    want = np.array([om.find_cell(p) for p in q])
    assert np.array_equal(got, want), np.nonzero(got != want)
    assert np.all(got[:-3] >= 0) and np.all(got[-3:] == -1)
    assert all(ctx.find_cell(p) == w for p, w in zip(q[:50], want[:50]))


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_row_block_spmm_matches_lane_spmm(precision, oracle_mod):
    """The LDS-staged row-block smoother SpMM (plain and persistent-pipelined), the LDS
    x-block one and the lanes-per-row one give the same solve (all to the exact
    discrete solution)."""
    m = mg.halfspace_mt(n=4, n_rx=3)
    sols = {}
    for rows in ("lanes", "rowblock", "pipelined", "xblock"):
        ctx = _ctx(m, "-inner_precision %s -spmm_rows %s" % (precision, rows))
        ctx.create_linear_system(2)
        ctx.assemble_matrix(0)
        ctx.assemble_rhs_mt(0, api.XY_POLAR, 0)
        ctx.assemble_rhs_mt(0, api.YX_POLAR, 1)
        ctx.create_pc()
        its, rr = ctx.solve_linear_system(0, 2, 300, 1e-10)
        assert np.all(rr <= 1e-10 * 1.0001)
        sols[rows] = [ctx.get_solution(s) for s in range(2)]
        rp, ci, C, M = ctx.get_matrix()
        for s in range(2):
            e = oracle_mod.solve_direct(rp, ci, C, M, *ctx.get_rhs(s))
            assert np.linalg.norm(sols[rows][s] - e) <= 1e4 * 1e-10 * np.linalg.norm(e)


@pytest.mark.parametrize("pc", ["blockdiag", "presb"])
@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_row_blocks_on_16bit_columns_match_32bit_columns(precision, pc):
    """The row-block SpMMs of B (smoother and fp64 FCG product) reading each column as
    a 16-bit offset from its 64-nonzero group's smallest column give the solve of the
    32-bit columns bit for bit, on 6 real columns (block-diagonal, XY + YX + dual) and
    on 3 (PRESB halves)."""
    m = mg.halfspace_mt(n=4, n_rx=3)
    out = {}
    for c16 in ("true", "false"):
        ctx = _ctx(m, "-inner_precision %s -spmm_rows rowblock -A_pc_type %s -spmm_col16 %s" % (precision, pc, c16))
        ctx.create_linear_system(3)
        ctx.assemble_matrix(0)
        ctx.assemble_rhs_mt(0, api.XY_POLAR, 0)
        ctx.assemble_rhs_mt(0, api.YX_POLAR, 1)
        ctx.assemble_rhs_dual(2)
        ctx.create_pc()
        assert ctx.stat("spmm_col16") == (1.0 if c16 == "true" else 0.0)
        its, rr = ctx.solve_linear_system(0, 3, 300, 1e-8)
        assert np.all(rr <= 1e-8 * 1.0001), rr
        out[c16] = (its, [ctx.get_solution(s) for s in range(3)])
        ctx.close()
    assert np.array_equal(out["true"][0], out["false"][0])
    for a, b in zip(out["true"][1], out["false"][1]):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_refreshed_preconditioner_matches_a_fresh_one(precision):
    """A second preconditioner set-up on the same linear system (the next frequency)
    refreshes the device state of the first (values of B, l1 scaling and the AMG
    hierarchy; buffers, transfers and tuned row mappings kept): its solve equals, bit
    for bit, that of a context set up at that frequency only.  A changed set-up
    option builds a new preconditioner instead."""
    m = mg.halfspace_mt(n=4, n_rx=3)
    m.freqs = np.array([0.3, 3.0])
    opts = "-inner_precision %s -spmm_rows rowblock" % precision

    def solve(ctx, f):
        ctx.assemble_matrix(f)
        ctx.assemble_rhs_mt(f, api.XY_POLAR, 0)
        ctx.assemble_rhs_mt(f, api.YX_POLAR, 1)
        ctx.assemble_rhs_dual(2)
        ctx.create_pc()
        its, rr = ctx.solve_linear_system(0, 3, 300, 1e-10)
        assert np.all(rr <= 1e-10 * 1.0001), rr
        return its, [ctx.get_solution(s) for s in range(3)]

    a = _ctx(m, opts)
    a.create_linear_system(3)
    solve(a, 0)
    assert a.stat("pc_reused") == 0.0
    its_a, sol_a = solve(a, 1)
    assert a.stat("pc_reused") == 1.0
    b = _ctx(m, opts)
    b.create_linear_system(3)
    its_b, sol_b = solve(b, 1)
    assert b.stat("pc_reused") == 0.0
    assert np.array_equal(its_a, its_b)
    for x, y in zip(sol_a, sol_b):
        assert np.array_equal(x, y)
    a.process_options("-ams_cheb_degree 5")
    solve(a, 0)
    assert a.stat("pc_reused") == 0.0
    a.close()
    b.close()


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_prepared_setup_matches_setup_in_turn(precision):
    """emfem_prepare_pc(next) started before the current solve: the next frequency's
    matrix (assembled on the context's second stream), its MT right-hand sides and its
    auxiliary-space set-up (built on worker threads) and so its solve equal, bit for
    bit, those of the same calls without it; a prepared frequency that is not the one
    assembled next is discarded and changes nothing."""
    m = mg.halfspace_mt(n=4, n_rx=3)
    m.freqs = np.array([0.3, 3.0, 30.0])
    opts = "-inner_precision %s -spmm_rows rowblock" % precision

    def step(ctx, f, nxt=None):
        ctx.assemble_matrix(f)
        ctx.assemble_rhs_mt(f, api.XY_POLAR, 0)
        ctx.assemble_rhs_mt(f, api.YX_POLAR, 1)
        ctx.assemble_rhs_dual(2)
        ctx.create_pc()
        if nxt is not None:
            ctx.prepare_pc(nxt)
        its, rr = ctx.solve_linear_system(0, 3, 300, 1e-10)
        assert np.all(rr <= 1e-10 * 1.0001), rr
        _, _, C, M = ctx.get_matrix()
        rhs = [np.concatenate(ctx.get_rhs(s)) for s in range(3)]
        return its, [ctx.get_solution(s) for s in range(3)], C, M, rhs

    a = _ctx(m, opts)
    a.create_linear_system(3)
    b = _ctx(m, opts)
    b.create_linear_system(3)
    step(a, 0, nxt=1)
    assert a.stat("pc_prepare_started") == 1.0
    step(b, 0)
    ra = step(a, 1, nxt=0)
    assert a.stat("assemble_prepared") == 1.0 and a.stat("pc_prepared") == 1.0
    assert a.stat("rhs_prepared") == 1.0
    rb = step(b, 1)
    assert b.stat("assemble_prepared") == 0.0 and b.stat("pc_prepared") == 0.0
    assert b.stat("rhs_prepared") == 0.0
    for x, y in zip(ra[2:4], rb[2:4]):
        assert np.array_equal(x, y)
    for x, y in zip(ra[4], rb[4]):
        assert np.array_equal(x, y)
    assert np.array_equal(ra[0], rb[0])
    for x, y in zip(ra[1], rb[1]):
        assert np.array_equal(x, y)
    # frequency 0 was prepared; assembling 2 instead discards it
    ra = step(a, 2)
    assert a.stat("assemble_prepared") == 0.0 and a.stat("pc_prepared") == 0.0
    rb = step(b, 2)
    assert np.array_equal(ra[0], rb[0])
    for x, y in zip(ra[1], rb[1]):
        assert np.array_equal(x, y)
    a.close()
    b.close()


@pytest.mark.parametrize("max_coarse,precision,pc", [(64, "fp32", "blockdiag"), (2048, "fp32", "blockdiag"),
                                                     (2048, "fp32", "presb"), (2048, "fp64", "presb")])
def test_coarse_size_reaches_the_exact_solution(max_coarse, precision, pc, oracle_mod):
    """The auxiliary AMG down to 64 unknowns (a deep hierarchy) or stopped at 2048 (one
    level, then the dense inverse of four components of a few hundred rows, read per
    64-row block over its components' columns only, for 3 and 6 columns in groups):
    each reaches the exact discrete solution to the bound its true residual sets.  At
    the same 1e-10 true residual the PRESB iterates keep more error on this mesh at
    every coarse size (2-7e-7 at 64, 256 and 2048, against 1-6e-8 block-diagonal),
    hence its looser bound."""
    m = mg.halfspace_mt(n=6, rho=100.0, freq=1.0, n_rx=2)
    ctx = _ctx(m, "-inner_precision %s -ams_max_coarse %d -A_pc_type %s" % (precision, max_coarse, pc))
    ctx.create_linear_system(3)
    ctx.assemble_matrix(0)
    ctx.assemble_rhs_mt(0, api.XY_POLAR, 0)
    ctx.assemble_rhs_mt(0, api.YX_POLAR, 1)
    ctx.assemble_rhs_dual(2)
    ctx.create_pc()
    rtol = 1e-10
    its, rr = ctx.solve_linear_system(0, 3, 300, rtol)
    print(max_coarse, precision, pc, "outer its", its, "true relres", rr, "inner its", ctx.stat("inner_its"))
    assert np.all(rr <= rtol * 1.0001), rr
    rp, ci, C, M = ctx.get_matrix()
    for slot in range(3):
        sr, si = ctx.get_rhs(slot)
        e = oracle_mod.solve_direct(rp, ci, C, M, sr, si)
        bound = (1e3 if pc == "blockdiag" else 1e4) * rtol
        assert np.linalg.norm(ctx.get_solution(slot) - e) <= bound * np.linalg.norm(e)
    ctx.close()


def test_coarse_size_out_of_range_is_an_argument_error():
    m = mg.halfspace_mt(n=3, n_rx=2)
    for bad in ("0", "4097"):
        ctx = _ctx(m, "-ams_max_coarse %s" % bad)
        ctx.create_linear_system(1)
        ctx.assemble_matrix(0)
        with pytest.raises(api.EmfemError):
            ctx.create_pc()
        ctx.close()


@pytest.mark.parametrize("cells,fits", [((64, 64, 8), True), ((72, 72, 8), False)])
def test_16bit_columns_only_where_every_group_fits(cells, fits):
    """The 16-bit column form is taken when the columns of every aligned group of 64
    nonzeros span at most 65536 (largest span 59,015 on 64 x 64 x 8 cells, 244,872
    edges) and the 32-bit columns are kept otherwise (74,455 on 72 x 72 x 8 cells,
    309,464 edges; spans computed from the oracle's pattern)."""
    m = mg.adaptive_mt(cells=cells, n_freqs=1, n_rx_side=2)
    ctx = _ctx(m, "-inner_precision fp32")
    ctx.create_linear_system(1)
    ctx.assemble_matrix(0)
    ctx.create_pc()
    assert ctx.stat("spmm_col16") == (1.0 if fits else 0.0)
    ctx.close()


@pytest.mark.parametrize("rows", ["auto", "lanes", "xblock"])
@pytest.mark.parametrize("n_tx", [10, 40])
def test_many_rhs_csem_block_solve(n_tx, rows, oracle_mod):
    """CSEM right-hand sides as one block (nc = 2 n_tx >= 16 real columns: the
    many-column SpMM kernels, tuned, lanes on columns only, or LDS x-blocks) reach
    the exact discrete solutions."""
    m = mg.tiny_mt(n=3, aniso_form=0, seed=1, freq=1.25)
    rng = np.random.default_rng(n_tx)
    lo, hi = m.points.min(0), m.points.max(0)
    tx = np.zeros((n_tx, 7))
    tx[:, 0] = rng.uniform(-2000.0, 2000.0, n_tx)
    tx[:, 1] = rng.uniform(-2000.0, 2000.0, n_tx)
    tx[:, 2] = rng.uniform(100.0, 1500.0, n_tx)
    tx[:, 3] = rng.uniform(0.0, 360.0, n_tx)
    tx[:, 5] = 1.0
    m.tx = tx
    ctx = _ctx(m, "-inner_precision fp32 -spmm_rows %s" % rows)
    ctx.create_linear_system(n_tx)
    ctx.assemble_matrix(0)
    for t in range(n_tx):
        ctx.assemble_rhs_csem(0, t, t)
    ctx.create_pc()
    its, rr = ctx.solve_linear_system(0, n_tx, 300, 1e-10)
    assert np.all(rr <= 1e-10 * 1.0001), rr
    rp, ci, C, M = ctx.get_matrix()
    for t in range(0, n_tx, 7):
        e = oracle_mod.solve_direct(rp, ci, C, M, *ctx.get_rhs(t))
        assert np.linalg.norm(ctx.get_solution(t) - e) <= 1e4 * 1e-10 * np.linalg.norm(e)
