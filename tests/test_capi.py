"""CPU tests of the C-ABI library: it loads, exports every declared symbol, and its
host-side logic (options, topology, point location, marking, error codes) agrees
with the oracle.  No compute kernels run here."""
import ctypes
import os
import re

import numpy as np
import pytest

from emfem_amd import api
from emfem_amd import meshgen as mg

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            names |= set(re.findall(r"\b(emfem_[a-z_0-9]+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(api.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 35
    for n in names:
        assert hasattr(lib, n), n
    # and the Python binding declares each of them
    assert set(names) == set(api.SIGNATURES), set(names) ^ set(api.SIGNATURES)


def test_options_parsing():
    ctx = api.Context()
    text = """# EMFEM options
-iprefix model   # trailing comment
-oprefix out
-inner_pc_type ams
-e_rtol 1e-9
-max_tx_edge_length -1
-K_max_it 50
-save_mesh
-A_ksp_gmres_restart 40
-B_pc_hypre_ams_cycle_type 13
"""
    ctx.process_options(text)
    assert ctx.get_option("iprefix") == "model"
    assert ctx.get_option("e_rtol") == "1e-9"
    assert ctx.get_option("max_tx_edge_length") == "-1"
    assert ctx.get_option("K_max_it") == "50"
    assert ctx.get_option("save_mesh") == "true"
    assert ctx.get_option("A_ksp_gmres_restart") == "40"
    assert ctx.get_option("dual_rtol") == "1e-6"         # default (src/em_ctx.cc:148-149)
    assert ctx.get_option("pc_threshold") == "500000"    # default (src/em_ctx.cc:136-137)
    assert ctx.unused_options() == ["B_pc_hypre_ams_cycle_type"]
    assert api.Context().get_option("inner_pc_type") == "ams"
    # the direct inner solver (src/em_la.cc:147-155) is not implemented: its options
    # are parsed and reported as not honoured
    for text, unused in (("-inner_pc_type mixed", ["inner_pc_type"]),
                         ("-inner_pc_type direct -direct_solver_type mumps",
                          ["direct_solver_type", "inner_pc_type"]),
                         ("-pc_threshold 1000", ["pc_threshold"]),
                         ("-inner_pc_type ams", [])):
        c2 = api.Context(text)
        assert sorted(c2.unused_options()) == unused, text
        c2.close()
    with pytest.raises(api.EmfemError) as e:
        ctx.process_options("-inner_pc_type bogus")
    assert e.value.code == api.ERR_USER
    with pytest.raises(api.EmfemError):
        ctx.get_option("no_such_key")


@pytest.fixture(scope="module")
def tiny_pair(oracle_mod):
    m = mg.tiny_mt(n=3, aniso_form=3, seed=1)
    ctx = api.Context()
    ctx.load_model(m)
    om = oracle_mod.Mesh(m.points, m.tets)
    return m, ctx, om


def test_topology_matches_oracle(tiny_pair):
    m, ctx, om = tiny_pair
    el, te, nb, eb = ctx.topology()
    assert np.array_equal(el, om.edgelist)
    assert np.array_equal(te, om.tet2edge)
    assert np.array_equal(nb, om.neighbor)
    assert np.array_equal(eb, om.edge_bdr)
    info = ctx.mesh_info()
    assert info["n_colours"] >= 1


def test_topology_plain_mt_counts(oracle_mod):
    m = mg.plain_mt()
    ctx = api.Context()
    ctx.load_model(m)
    info = ctx.mesh_info()
    assert (info["n_vertices"], info["n_tets"], info["n_edges"]) == (9702, 50400, 62581)


def test_user_edge_numbering(tiny_pair, oracle_mod):
    """TetGen-style supplied edges (arbitrary numbering and local order) are honoured."""
    m, _, om = tiny_pair
    rng = np.random.default_rng(2)
    perm = rng.permutation(om.n_edges)
    el = om.edgelist[perm][:, ::-1].copy()          # renumbered, endpoints reversed
    inv = np.empty_like(perm)
    inv[perm] = np.arange(perm.size)
    t2e = inv[om.tet2edge]
    t2e = np.array([row[rng.permutation(6)] for row in t2e], dtype=np.int32)
    ctx = api.Context()
    ctx.load_model(m)
    ctx.set_mesh_edges(el, t2e)
    el2, te2, _, eb2 = ctx.topology()
    assert np.array_equal(el2, np.sort(el, axis=1))
    assert np.array_equal(np.sort(te2, axis=1), np.sort(t2e, axis=1))
    assert np.array_equal(eb2, om.edge_bdr[perm])
    bad = t2e.copy()
    bad[0, 0] = bad[0, 1]
    with pytest.raises(api.EmfemError):
        ctx.set_mesh_edges(el, bad)


def test_point_location_matches_oracle(tiny_pair):
    m, ctx, om = tiny_pair
    rng = np.random.default_rng(3)
    pts = np.concatenate([rng.uniform(om.points.min(0), om.points.max(0), (40, 3)), om.points[::7], m.rx])
    for p in pts:
        assert ctx.find_cell(p) == om.find_cell(p)
    assert ctx.find_cell(om.points.max(0) + 10.0) == -1


def test_background_model_host_only(tiny_pair):
    _, ctx, _ = tiny_pair
    ctx.update_background_model()


def test_errors_and_device_requirement():
    m = mg.tiny_mt(n=2)
    ctx = api.Context()
    with pytest.raises(api.EmfemError) as e:
        ctx.set_rho(7, np.ones((3, 1)))
    assert e.value.code == api.ERR_USER and "anisotropy" in str(e.value)
    with pytest.raises(api.EmfemError) as e:
        ctx.assemble_matrix(0)
    assert e.value.code == api.ERR_STATE
    ctx.load_model(m)
    bad = m.attr.copy()
    bad[0] = 99
    with pytest.raises(api.EmfemError):
        ctx.set_mesh(m.points, m.tets, bad)
    ctx.set_mesh(m.points, m.tets, m.attr)
    try:
        ctx.device_array(np.zeros(1)).free()
        has_gpu = True
    except api.EmfemError:
        has_gpu = False
    if not has_gpu:
        with pytest.raises(api.EmfemError) as e:
            ctx.create_linear_system(1)
        assert e.value.code == api.ERR_HIP


def test_marking_matches_oracle(oracle_mod):
    rng = np.random.default_rng(4)
    for _ in range(10):
        err = rng.exponential(size=rng.integers(3, 300))
        frac = rng.uniform(0.01, 0.95)
        assert np.array_equal(api.refine_fixed_fraction(err, frac), oracle_mod.refine_fixed_fraction(err, frac))
        assert np.array_equal(api.refine_fixed_number(err, frac), oracle_mod.refine_fixed_number(err, frac))
