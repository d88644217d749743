"""The allocation-free split-layout operator (emfem_matshell_mult_a_split_noalloc)
against emfem_matshell_mult_a_split, through the C ABI: bitwise equal results on two
calls in a row, the second with a smaller input (fewer right-hand sides, then a
smaller mesh), and a clear error above its staging limit."""
import numpy as np
import pytest

from emfem_amd import api
from emfem_amd import meshgen as mg

pytestmark = pytest.mark.gpu

LIMIT = 2097152  # n_edges * n_rhs the fixed staging array holds


def _system(m):
    c = api.Context()
    c.load_model(m)
    c.create_linear_system(1)
    c.assemble_matrix(0)
    return c


def _check(c, n_rhs, rng):
    ne = c.mesh_info()["n_edges"]
    xr, xi = c.create_vecs(n_rhs, api.LAYOUT_SPLIT)
    ar, ai = c.create_vecs(n_rhs, api.LAYOUT_SPLIT)
    br, bi = c.create_vecs(n_rhs, api.LAYOUT_SPLIT)
    try:
        c.copy_to_device(xr, rng.standard_normal(ne * n_rhs))
        c.copy_to_device(xi, rng.standard_normal(ne * n_rhs))
        c.matshell_mult_a_split(n_rhs, xr, xi, ar, ai)
        c.matshell_mult_a_split_noalloc(n_rhs, xr, xi, br, bi)
        for u, v in ((ar, br), (ai, bi)):
            a = c.copy_to_host(u, np.empty(ne * n_rhs))
            b = c.copy_to_host(v, np.empty(ne * n_rhs))
            assert np.array_equal(a, b)
            assert np.any(a != 0)
    finally:
        for p in (xr, xi, ar, ai, br, bi):
            c.device_free(p)


def test_noalloc_operator_matches_split_operator():
    rng = np.random.default_rng(5)
    big = _system(mg.tiny_mt(n=4, aniso_form=3, seed=2, freq=0.9))
    _check(big, 5, rng)
    _check(big, 2, rng)          # second call, smaller input
    small = _system(mg.tiny_mt(n=2, aniso_form=1, seed=3, freq=2.0))
    _check(small, 1, rng)        # smaller mesh on the same GPU
    _check(big, 3, rng)
    ne = big.mesh_info()["n_edges"]
    n_over = LIMIT // ne + 1
    with pytest.raises(api.EmfemError) as e:
        big.matshell_mult_a_split_noalloc(n_over, 1, 1, 1, 1)   # rejected before any device access
    assert e.value.code == api.ERR_ARG and "exceeds" in str(e.value)
    big.close()
    small.close()
