"""Python face of the CPU FP64 oracle (test infrastructure only).

Wraps `oracle/liboracle.so` (the C restatement in `oracle.c`) with numpy arrays and adds
the two pieces that are plain linear algebra:

* `apply_operator` restates `matshell_mult_a` (src/em_la.cc:65-87) with scipy.sparse;
* `solve_direct` gives the exact discrete solution of the real-equivalent system that
  `solve_linear_system` (src/em_la.cc:35-63) converges to, by a sparse LU of
  K = C + iM applied to s_re - i s_im.  The reference's own `direct` inner solver is a
  Cholesky of B (src/em_la.cc:147-155); any exact solver reaches the same solution.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg import this module.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
MAXLAY = 256

_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", HERE, "-B"])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


P = ctypes.c_void_p
I = ctypes.c_int
D = ctypes.c_double


def _declare(L):
    sig = {
        "or_build_edges": (I, [I, P, P, P]),
        "or_neighbors": (None, [I, P, P, P]),
        "or_boundary_markers": (None, [I, P, P, P, P, P, P]),
        "or_make_pattern": (I, [I, I, P, P, P, I]),
        "or_sigma": (I, [I, P, P]),
        "or_element_matrices": (I, [P, P, P, P, I, P, D, P, P]),
        "or_tet_geometry": (None, [P, P, I, P, P]),
        "or_assemble_matrix": (I, [I, P, P, P, P, P, P, I, P, D, P, P, P, P, P]),
        "or_assemble_matrix_mt": (I, [I, P, P, P, P, P, P, I, P, D, P, P, P, P, P, I]),
        "or_estimate_error_mt": (I, [I, P, P, P, P, P, P, P, I, P, D, I, P, P, I]),
        "or_contains_point": (I, [P, P, I, P]),
        "or_vertex_to_tet": (None, [I, P, P]),
        "or_nearest_vertex": (I, [I, P, P]),
        "or_find_cell": (I, [I, P, I, P, P, P, P]),
        "or_pw_eval": (None, [I, P, P, D, D, I, P, P]),
        "or_background_model": (I, [I, P, I, P, P, P, I, P, P, P, P, P]),
        "or_blend_weights": (None, [P, P, P]),
        "or_boundary_values": (I, [I, P, P, P, P, P, P, P, P, P, D, I, I, P, P]),
        "or_assemble_rhs_mt": (I, [I, P, P, P, P, P, P, I, P, D, P, P, P]),
        "or_divide_line_source": (I, [P, I, P]),
        "or_assemble_rhs_csem": (I, [I, P, I, P, P, P, P, D, P, I, P, P]),
        "or_assemble_rhs_dual": (I, [I, P, I, P, P, P, P, I, P, P, P]),
        "or_interpolate_fields": (None, [P, P, P, P, I, P, D, P, P, P]),
        "or_mt_response": (I, [I, D, P, P, P, P, P]),
        "or_csem_response": (I, [I, P, P, P]),
        "or_tet_diameter": (D, [P, P, I]),
        "or_face_diameter": (D, [P, P, I, I]),
        "or_estimate_error": (I, [I, P, P, P, P, P, P, P, I, P, D, I, P, P]),
        "or_refine_fixed_fraction": (None, [I, P, D, P]),
        "or_refine_fixed_number": (None, [I, P, D, P]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


class Mesh:
    """Topology the reference derives from TetGen's output (src/em_mesh.cc:439-502)."""

    def __init__(self, points, tets, edgelist=None, tet2edge=None):
        L = lib()
        self.points = _c(points, np.float64)
        self.tets = _c(tets, np.int32)
        nt = self.tets.shape[0]
        if edgelist is None:
            el = np.empty((nt * 6, 2), dtype=np.int32)
            t2e = np.empty((nt, 6), dtype=np.int32)
            ne = L.or_build_edges(nt, _p(self.tets), _p(el), _p(t2e))
            self.edgelist = np.ascontiguousarray(el[:ne])
            self.tet2edge = t2e
        else:
            el = _c(edgelist, np.int32).copy()
            el.sort(axis=1)  # orientation: lower vertex id first (src/em_mesh.cc:448-455)
            self.edgelist = el
            self.tet2edge = _c(tet2edge, np.int32)
        self.neighbor = np.empty((nt, 4), dtype=np.int32)
        self.non = np.empty((nt, 4), dtype=np.int32)
        L.or_neighbors(nt, _p(self.tets), _p(self.neighbor), _p(self.non))
        self.edge_bdr = np.zeros(self.n_edges, dtype=np.uint8)
        self.vertex_bdr = np.zeros(self.n_vertices, dtype=np.uint8)
        L.or_boundary_markers(nt, _p(self.tets), _p(self.neighbor), _p(self.tet2edge), _p(self.edgelist),
                              _p(self.edge_bdr), _p(self.vertex_bdr))
        self.v2t = np.empty(self.n_vertices, dtype=np.int32)
        L.or_vertex_to_tet(nt, _p(self.tets), _p(self.v2t))

    @property
    def n_tets(self):
        return self.tets.shape[0]

    @property
    def n_vertices(self):
        return self.points.shape[0]

    @property
    def n_edges(self):
        return self.edgelist.shape[0]

    def pattern(self):
        """CSR pattern (src/em_fwd.cc:191-246)."""
        ne, nt = self.n_edges, self.n_tets
        rowptr = np.empty(ne + 1, dtype=np.int32)
        cap = nt * 36
        col = np.empty(cap, dtype=np.int32)
        nnz = lib().or_make_pattern(ne, nt, _p(self.tet2edge), _p(rowptr), _p(col), cap)
        assert nnz >= 0
        return rowptr, np.ascontiguousarray(col[:nnz])

    def find_cell(self, p):
        p = _c(p, np.float64)
        return lib().or_find_cell(self.n_vertices, _p(self.points), self.n_tets, _p(self.tets),
                                  _p(self.neighbor), _p(self.v2t), _p(p))

    def contains(self, t, p):
        p = _c(p, np.float64)
        return bool(lib().or_contains_point(_p(self.points), _p(self.tets), int(t), _p(p)))

    def gradient_matrix(self):
        """Discrete gradient G: row e = -1 at the lower vertex, +1 at the higher one
        (src/em_la.cc:254-277)."""
        import scipy.sparse as sp
        ne = self.n_edges
        rows = np.repeat(np.arange(ne), 2)
        cols = self.edgelist.ravel()
        vals = np.tile([-1.0, 1.0], ne)
        return sp.csr_matrix((vals, (rows, cols)), shape=(ne, self.n_vertices))


def sigma(aniso_form, row):
    out = np.empty(9)
    rc = lib().or_sigma(int(aniso_form), _p(_c(row, np.float64)), _p(out))
    if rc != 0:
        raise ValueError("unsupported anisotropy form %d" % aniso_form)
    return out.reshape(3, 3)


def element_matrices(mesh, t, sig, omega):
    C = np.empty(36)
    M = np.empty(36)
    lib().or_element_matrices(_p(mesh.points), _p(mesh.tets), _p(mesh.tet2edge), _p(mesh.edgelist), int(t),
                              _p(_c(sig, np.float64).ravel()), float(omega), _p(C), _p(M))
    return C.reshape(6, 6), M.reshape(6, 6)


def tet_geometry(mesh, t):
    coef = np.empty(16)
    det = np.empty(1)
    lib().or_tet_geometry(_p(mesh.points), _p(mesh.tets), int(t), _p(coef), _p(det))
    return coef.reshape(4, 4), float(det[0])


def assemble(mesh, model, freq, pattern=None, threads=1):
    """C, M values on the CSR pattern plus the bdr-cell flags (src/em_fwd.cc:318-400).
    threads > 1 runs the OpenMP loop (atomic scatter: equal to rounding, not bitwise)."""
    rowptr, col = pattern if pattern is not None else mesh.pattern()
    C = np.zeros(col.size)
    M = np.zeros(col.size)
    bdr = np.zeros(mesh.n_tets, dtype=np.uint8)
    rho = _c(model.rho, np.float64)
    attr = _c(model.attr, np.int32)
    args = (mesh.n_tets, _p(mesh.points), _p(mesh.tets), _p(mesh.tet2edge), _p(mesh.edgelist), _p(mesh.edge_bdr),
            _p(attr), int(model.aniso_form), _p(rho), float(freq), _p(rowptr), _p(col), _p(C), _p(M), _p(bdr))
    if threads > 1:
        rc = lib().or_assemble_matrix_mt(*args, int(threads))
    else:
        rc = lib().or_assemble_matrix(*args)
    if rc != 0:
        raise ValueError("assembly failed")
    return rowptr, col, C, M, bdr


class Background:
    def __init__(self, corners, nlay, ztop, lsig):
        self.corners, self.nlay, self.ztop, self.lsig = corners, nlay, ztop, lsig

    def layers(self, k):
        n = self.nlay[k]
        return self.ztop[k, :n].copy(), self.lsig[k, :n].copy()


def background_model(mesh, model):
    """update_background_model (src/em_fwd.cc:135-159)."""
    corners = np.empty((4, 3))
    nlay = np.zeros(4, dtype=np.int32)
    ztop = np.zeros((4, MAXLAY))
    lsig = np.zeros((4, MAXLAY))
    rc = lib().or_background_model(mesh.n_vertices, _p(mesh.points), mesh.n_tets, _p(mesh.tets),
                                   _p(mesh.neighbor), _p(_c(model.attr, np.int32)), int(model.aniso_form),
                                   _p(_c(model.rho, np.float64)), _p(corners), _p(nlay), _p(ztop), _p(lsig))
    if rc != 0:
        raise ValueError("background model extraction failed")
    return Background(corners, nlay, ztop, lsig)


def pw_eval(ztop, lsig, freq, z, mode):
    e = np.empty(6)
    h = np.empty(6)
    lib().or_pw_eval(len(ztop), _p(_c(ztop, np.float64)), _p(_c(lsig, np.float64)), float(freq), float(z),
                     int(mode), _p(e), _p(h))
    return e[0::2] + 1j * e[1::2], h[0::2] + 1j * h[1::2]


def blend_weights(corners, ec):
    """Weights of the four corner columns at a boundary-edge midpoint (src/em_fwd.cc:438-471)."""
    out = np.empty(4)
    lib().or_blend_weights(_p(_c(corners, np.float64)), _p(_c(ec, np.float64)), _p(out))
    return out


def boundary_values(mesh, bg, freq, mode):
    """Dirichlet DOF values on boundary edges (src/em_fwd.cc:402-486) -> complex[E] + mask."""
    bval = np.zeros((mesh.n_edges, 2))
    have = np.zeros(mesh.n_edges, dtype=np.uint8)
    lib().or_boundary_values(mesh.n_tets, _p(mesh.points), _p(mesh.tets), _p(mesh.tet2edge), _p(mesh.edgelist),
                             _p(mesh.edge_bdr), _p(bg.corners), _p(bg.nlay), _p(bg.ztop), _p(bg.lsig),
                             float(freq), int(mode), mesh.n_edges, _p(bval), _p(have))
    return bval[:, 0] + 1j * bval[:, 1], have.astype(bool)


def rhs_mt(mesh, model, bg, freq, mode):
    """assemble_rhs_mt (src/em_fwd.cc:488-580) -> (s_re, s_im)."""
    bv, _ = boundary_values(mesh, bg, freq, mode)
    bval = np.ascontiguousarray(np.stack([bv.real, bv.imag], axis=1))
    s_re = np.zeros(mesh.n_edges)
    s_im = np.zeros(mesh.n_edges)
    rc = lib().or_assemble_rhs_mt(mesh.n_tets, _p(mesh.points), _p(mesh.tets), _p(mesh.tet2edge),
                                  _p(mesh.edgelist), _p(mesh.edge_bdr), _p(_c(model.attr, np.int32)),
                                  int(model.aniso_form), _p(_c(model.rho, np.float64)), float(freq), _p(bval),
                                  _p(s_re), _p(s_im))
    if rc != 0:
        raise ValueError("rhs_mt failed")
    return s_re, s_im


def divide_line_source(tx, n_div):
    out = np.empty((max(1, n_div), 8))
    n = lib().or_divide_line_source(_p(_c(tx, np.float64)), int(n_div), _p(out))
    return out[:n]


def rhs_csem(mesh, freq, tx, n_div=20):
    """assemble_rhs_csem (src/em_fwd.cc:616-671) -> (s_re, s_im)."""
    s_re = np.zeros(mesh.n_edges)
    s_im = np.zeros(mesh.n_edges)
    rc = lib().or_assemble_rhs_csem(mesh.n_vertices, _p(mesh.points), mesh.n_tets, _p(mesh.tets),
                                    _p(mesh.tet2edge), _p(mesh.edgelist), _p(mesh.neighbor), float(freq),
                                    _p(_c(tx, np.float64)), int(n_div), _p(s_re), _p(s_im))
    if rc != 0:
        raise ValueError("transmitter not found in mesh")
    return s_re, s_im


def rhs_dual(mesh, rx):
    """assemble_rhs_dual (src/em_fwd.cc:673-717) -> (s_re, s_im)."""
    rx = _c(rx, np.float64)
    s_re = np.zeros(mesh.n_edges)
    s_im = np.zeros(mesh.n_edges)
    rc = lib().or_assemble_rhs_dual(mesh.n_vertices, _p(mesh.points), mesh.n_tets, _p(mesh.tets),
                                    _p(mesh.tet2edge), _p(mesh.edgelist), _p(mesh.neighbor), rx.shape[0],
                                    _p(rx), _p(s_re), _p(s_im))
    if rc != 0:
        raise ValueError("receiver not found in mesh")
    return s_re, s_im


def csr(rowptr, col, val, n):
    import scipy.sparse as sp
    return sp.csr_matrix((val, col, rowptr), shape=(n, n))


def apply_operator(rowptr, col, C, M, xr, xi):
    """matshell_mult_a: y_r = C x_r - M x_i ; y_i = -M x_r - C x_i (src/em_la.cc:65-87)."""
    n = rowptr.size - 1
    Cm, Mm = csr(rowptr, col, C, n), csr(rowptr, col, M, n)
    return Cm @ xr - Mm @ xi, -(Mm @ xr) - Cm @ xi


def direct_solver(rowptr, col, C, M):
    """Factorises K = C + iM once; returns solve(s_re, s_im) -> e for
    A [e_re; e_im] = [s_re; s_im], i.e. (C + iM) e = s_re - i s_im
    (src/em_la.cc:35-87).  K is complex symmetric with a symmetric pattern, so the
    sparse LU runs in symmetric mode (diagonal pivots, minimum-degree ordering of
    K + K^T), which has far less fill than the default; every solve is checked by
    its residual, and a factorisation with threshold pivoting replaces the symmetric
    one if that check fails."""
    import scipy.sparse.linalg as spla
    n = rowptr.size - 1
    K = csr(rowptr, col, C.astype(np.complex128) + 1j * M, n).tocsc()
    state = {"lu": spla.splu(K, permc_spec="MMD_AT_PLUS_A", diag_pivot_thresh=0.0,
                             options=dict(SymmetricMode=True)), "pivoting": False}

    def solve(s_re, s_im):
        rhs = (np.asarray(s_re) - 1j * np.asarray(s_im)).astype(np.complex128)
        x = state["lu"].solve(rhs)
        nb = np.linalg.norm(rhs)
        if nb > 0 and np.linalg.norm(K @ x - rhs) > 1e-11 * nb and not state["pivoting"]:
            state["lu"] = spla.splu(K, permc_spec="MMD_AT_PLUS_A")
            state["pivoting"] = True
            x = state["lu"].solve(rhs)
        return x

    return solve


def solve_direct(rowptr, col, C, M, s_re, s_im):
    """Exact solution e = e_re + i e_im of A [e_re; e_im] = [s_re; s_im], i.e.
    (C + iM) e = s_re - i s_im (src/em_la.cc:35-87)."""
    return direct_solver(rowptr, col, C, M)(s_re, s_im)


def interpolate_fields(mesh, t, p, freq, e):
    """E and H at p inside tet t (src/em_fwd.cc:719-744); e is complex[E]."""
    ev = np.ascontiguousarray(np.stack([e.real, e.imag], axis=1))
    es = np.empty(6)
    hs = np.empty(6)
    lib().or_interpolate_fields(_p(mesh.points), _p(mesh.tets), _p(mesh.tet2edge), _p(mesh.edgelist), int(t),
                                _p(_c(p, np.float64)), float(freq), _p(ev), _p(es), _p(hs))
    return es[0::2] + 1j * es[1::2], hs[0::2] + 1j * hs[1::2]


def mt_response(otype, freq, exy, hxy, eyx, hyx):
    def ri(z):
        return np.ascontiguousarray(np.stack([np.real(z), np.imag(z)], axis=1).ravel())
    out = np.empty(2)
    rc = lib().or_mt_response(int(otype), float(freq), _p(ri(exy)), _p(ri(hxy)), _p(ri(eyx)), _p(ri(hyx)),
                              _p(out))
    if rc != 0:
        raise ValueError("data type %d not implemented" % otype)
    return out[0] + 1j * out[1]


def csem_response(otype, es, hs):
    def ri(z):
        return np.ascontiguousarray(np.stack([np.real(z), np.imag(z)], axis=1).ravel())
    out = np.empty(2)
    rc = lib().or_csem_response(int(otype), _p(ri(es)), _p(ri(hs)), _p(out))
    if rc != 0:
        raise ValueError("data type %d not implemented" % otype)
    return out[0] + 1j * out[1]


def responses_mt(mesh, model, fidx, e_xy, e_yx):
    """calculate_response_mt (src/em_fwd.cc:746-802): rsp[o] for MT observations of fidx."""
    out = {}
    freq = float(model.freqs[fidx])
    cache = {}
    for o, (ot, fi, ti, ri) in enumerate(model.obs):
        if fi != fidx or ti != -3:
            continue
        if ri not in cache:
            t = mesh.find_cell(model.rx[ri])
            if t < 0:
                raise ValueError("receiver %d is not found in mesh" % ri)
            exy, hxy = interpolate_fields(mesh, t, model.rx[ri], freq, e_xy)
            eyx, hyx = interpolate_fields(mesh, t, model.rx[ri], freq, e_yx)
            cache[ri] = (exy, hxy, eyx, hyx)
        out[o] = mt_response(ot, freq, *cache[ri])
    return out


def estimate_error(mesh, model, freq, sols, threads=1):
    """Per-tet squared indicators for each solution (src/em_fwd.cc:1005-1170);
    threads > 1 runs the OpenMP loop (bitwise the serial result)."""
    L = lib()
    nt = mesh.n_tets
    bufs = [np.ascontiguousarray(np.stack([s.real, s.imag], axis=1)) for s in sols]
    arr = (ctypes.c_void_p * len(bufs))(*[b.ctypes.data for b in bufs])
    errors = np.empty((len(sols), nt))
    args = (nt, _p(mesh.points), _p(mesh.tets), _p(mesh.tet2edge), _p(mesh.edgelist), _p(mesh.neighbor),
            _p(mesh.non), _p(_c(model.attr, np.int32)), int(model.aniso_form), _p(_c(model.rho, np.float64)),
            float(freq), len(sols), ctypes.cast(arr, P), _p(errors))
    if threads > 1:
        rc = L.or_estimate_error_mt(*args, int(threads))
    else:
        rc = L.or_estimate_error(*args)
    if rc != 0:
        raise ValueError("estimator failed")
    return errors


def combine_indicators(errors, mt=True):
    """sqrt(eta_primal * eta_dual) per polarisation (src/em_fwd.cc:1172-1177)."""
    if mt:
        return np.sqrt(errors[0] * errors[2]), np.sqrt(errors[1] * errors[2])
    return (np.sqrt(errors[0] * errors[1]),)


def refine_fixed_fraction(error, fraction, flag=None):
    error = _c(error, np.float64)
    flag = np.zeros(error.size, dtype=np.uint8) if flag is None else flag
    lib().or_refine_fixed_fraction(error.size, _p(error), float(fraction), _p(flag))
    return flag


def refine_fixed_number(error, fraction, flag=None):
    error = _c(error, np.float64)
    flag = np.zeros(error.size, dtype=np.uint8) if flag is None else flag
    lib().or_refine_fixed_number(error.size, _p(error), float(fraction), _p(flag))
    return flag
