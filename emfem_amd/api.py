"""ctypes binding of libemfem_amd.so (include/emfem_amd.h).

`Context` mirrors the reference's EMContext entry points one to one
(create_context / process_options / assemble_matrix / assemble_rhs_* / create_pc /
solve_linear_system / matshell_mult_a / pc_apply_b / estimate_error /
calculate_response_*; src/em_ctx.h:66-68, src/em_fwd.h:17-51, src/em_la.h:10-20).
Failures raise `EmfemError` carrying the status code and the library's message,
the way the reference propagates PetscErrorCode with its SETERRQ text.

There is no fallback: if the shared library is missing, importing this module
raises, and calls that need the GPU fail with EMFEM_ERR_HIP.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libemfem_amd.so")

OK, ERR_USER, ERR_HIP, ERR_NOT_CONVERGED, ERR_STATE, ERR_ARG = 0, -1, -2, -3, -4, -5
XY_POLAR, YX_POLAR, MT_TIDX = -1, -2, -3
LAYOUT_INTERLEAVED, LAYOUT_SPLIT = 0, 1


class EmfemError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("emfem error %d: %s" % (code, msg))
        self.code = code


P = ctypes.c_void_p
I = ctypes.c_int
D = ctypes.c_double
S = ctypes.c_char_p

# name -> (restype, argtypes); kept in the order of include/emfem_amd.h
SIGNATURES = {
    "emfem_create_context": (I, [ctypes.POINTER(P)]),
    "emfem_destroy_context": (I, [P]),
    "emfem_last_error": (S, [P]),
    "emfem_process_options": (I, [P, S]),
    "emfem_get_option": (I, [P, S, ctypes.c_char_p, I]),
    "emfem_unused_options": (I, [P, ctypes.c_char_p, I]),
    "emfem_set_mesh": (I, [P, I, P, I, P, P]),
    "emfem_set_mesh_edges": (I, [P, I, P, P]),
    "emfem_set_rho": (I, [P, I, I, P]),
    "emfem_set_frequencies": (I, [P, I, P]),
    "emfem_set_transmitters": (I, [P, I, P]),
    "emfem_set_receivers": (I, [P, I, P]),
    "emfem_get_mesh_info": (I, [P, P]),
    "emfem_get_topology": (I, [P, P, P, P, P]),
    "emfem_find_cell": (I, [P, P]),
    "emfem_update_background_model": (I, [P]),
    "emfem_create_linear_system": (I, [P, I]),
    "emfem_destroy_linear_system": (I, [P]),
    "emfem_assemble_matrix": (I, [P, I]),
    "emfem_get_matrix": (I, [P, P, P, P, P]),
    "emfem_assemble_rhs_mt": (I, [P, I, I, I]),
    "emfem_assemble_rhs_csem": (I, [P, I, I, I]),
    "emfem_assemble_rhs_dual": (I, [P, I]),
    "emfem_set_rhs": (I, [P, I, P, P]),
    "emfem_get_rhs": (I, [P, I, P, P]),
    "emfem_get_solution": (I, [P, I, P, P]),
    "emfem_create_pc": (I, [P]),
    "emfem_prepare_pc": (I, [P, I]),
    "emfem_destroy_pc": (I, [P]),
    "emfem_solve_linear_system": (I, [P, I, I, I, D, P, P, P]),
    "emfem_matshell_mult_a": (I, [P, I, P, P]),
    "emfem_pc_apply_b": (I, [P, I, P, P]),
    "emfem_create_vecs": (I, [P, I, I, ctypes.POINTER(P), ctypes.POINTER(P)]),
    "emfem_matshell_mult_a_split": (I, [P, I, P, P, P, P]),
    "emfem_pc_apply_b_split": (I, [P, I, P, P, P, P]),
    "emfem_matshell_mult_a_split_noalloc": (I, [P, I, P, P, P, P]),
    "emfem_estimate_error": (I, [P, I, I, P, P]),
    "emfem_interpolate_fields": (I, [P, I, I, I, P, P, P]),
    "emfem_calculate_response_mt": (I, [P, I, I, I, I, P, P]),
    "emfem_calculate_response_csem": (I, [P, I, I, I, I, P, P]),
    "emfem_refine_fixed_fraction": (I, [I, P, D, P]),
    "emfem_refine_fixed_number": (I, [I, P, D, P]),
    "emfem_read_mesh": (I, [P, S]),
    "emfem_read_emd": (I, [P, S]),
    "emfem_set_observations": (I, [P, I, P, P]),
    "emfem_save_rsp": (I, [P, S]),
    "emfem_get_responses": (I, [P, P]),
    "emfem_locate_points": (I, [P, I, P, P]),
    "emfem_solve_batch": (I, [P, I, I, I, I, D, P, P, P]),
    "emfem_surrogate_init": (I, [P, ctypes.c_uint]),
    "emfem_surrogate_params": (I, [P, P, P, P]),
    "emfem_surrogate_features": (I, [P, I, I, I, P]),
    "emfem_surrogate_train": (I, [P, I, P, P, P, I, D, P]),
    "emfem_surrogate_predict": (I, [P, I, I, I, P]),
    "emfem_surrogate_eval": (I, [P, I, P, P, P, P, P, P]),
    "emfem_save_mesh": (I, [P, S, I, P, P]),
    "emfem_forward": (I, [P]),
    "emfem_refine_mesh": (I, [P, P]),
    "emfem_refine_uniform": (I, [P]),
    "emfem_get_mesh": (I, [P, P, P, P]),
    "emfem_device_alloc": (I, [P, ctypes.c_ulonglong, ctypes.POINTER(P)]),
    "emfem_device_free": (I, [P, P]),
    "emfem_copy_to_device": (I, [P, P, P, ctypes.c_ulonglong]),
    "emfem_copy_to_host": (I, [P, P, P, ctypes.c_ulonglong]),
    "emfem_get_stat": (I, [P, S, ctypes.POINTER(D)]),
    "emfem_kernel_time": (I, [P, S, ctypes.POINTER(D), ctypes.POINTER(D)]),
}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libemfem_amd.so is not built (run __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


class Context:
    """One EMFEM context: a mesh, its linear system and preconditioner on one GPU."""

    def __init__(self, options: str | None = None):
        self._L = lib()
        h = P()
        rc = self._L.emfem_create_context(ctypes.byref(h))
        if rc != OK:
            raise EmfemError(rc, "create_context failed")
        self._h = h
        self._keep = []
        if options:
            self.process_options(options)

    # -- plumbing ---------------------------------------------------------------
    def _check(self, rc):
        if rc != OK:
            msg = self._L.emfem_last_error(self._h)
            raise EmfemError(rc, msg.decode() if msg else "")
        return rc

    def close(self):
        if getattr(self, "_h", None):
            self._L.emfem_destroy_context(self._h)
            self._h = None

    destroy_context = close

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- options ---------------------------------------------------------------
    def process_options(self, text: str):
        self._check(self._L.emfem_process_options(self._h, text.encode()))

    def get_option(self, key: str) -> str:
        buf = ctypes.create_string_buffer(256)
        self._check(self._L.emfem_get_option(self._h, key.encode(), buf, 256))
        return buf.value.decode()

    def unused_options(self):
        buf = ctypes.create_string_buffer(4096)
        self._check(self._L.emfem_unused_options(self._h, buf, 4096))
        return buf.value.decode().split()

    # -- model -----------------------------------------------------------------
    def set_mesh(self, points, tets, attributes):
        pts = _c(points, np.float64)
        tt = _c(tets, np.int32)
        at = _c(attributes, np.int32)
        self._check(self._L.emfem_set_mesh(self._h, pts.shape[0], _p(pts), tt.shape[0], _p(tt), _p(at)))

    def set_mesh_edges(self, edgelist, tet2edge):
        el = _c(edgelist, np.int32)
        te = _c(tet2edge, np.int32)
        self._check(self._L.emfem_set_mesh_edges(self._h, el.shape[0], _p(el), _p(te)))

    def set_rho(self, aniso_form, rho):
        r = _c(rho, np.float64)
        self._check(self._L.emfem_set_rho(self._h, r.shape[0], int(aniso_form), _p(r)))

    def set_frequencies(self, freqs):
        f = _c(freqs, np.float64)
        self._check(self._L.emfem_set_frequencies(self._h, f.size, _p(f)))

    def set_transmitters(self, tx):
        t = _c(np.asarray(tx).reshape(-1, 7), np.float64)
        self._check(self._L.emfem_set_transmitters(self._h, t.shape[0], _p(t)))

    def set_receivers(self, rx):
        r = _c(np.asarray(rx).reshape(-1, 3), np.float64)
        self._check(self._L.emfem_set_receivers(self._h, r.shape[0], _p(r)))

    def load_model(self, m):
        """read_mesh + read_rho + read_emd from a meshgen.Model."""
        self.set_mesh(m.points, m.tets, m.attr)
        self.set_rho(m.aniso_form, m.rho)
        self.set_frequencies(m.freqs)
        if len(m.tx):
            self.set_transmitters(m.tx)
        self.set_receivers(m.rx)

    def mesh_info(self):
        info = np.zeros(5, dtype=np.int32)
        self._check(self._L.emfem_get_mesh_info(self._h, _p(info)))
        return dict(n_vertices=int(info[0]), n_tets=int(info[1]), n_edges=int(info[2]), nnz=int(info[3]),
                    n_colours=int(info[4]))

    def topology(self):
        inf = self.mesh_info()
        el = np.empty((inf["n_edges"], 2), dtype=np.int32)
        te = np.empty((inf["n_tets"], 6), dtype=np.int32)
        nb = np.empty((inf["n_tets"], 4), dtype=np.int32)
        eb = np.empty(inf["n_edges"], dtype=np.uint8)
        self._check(self._L.emfem_get_topology(self._h, _p(el), _p(te), _p(nb), _p(eb)))
        return el, te, nb, eb

    def find_cell(self, p):
        q = _c(p, np.float64)
        return int(self._L.emfem_find_cell(self._h, _p(q)))

    def locate_points(self, points):
        """Batch point location on the device (find_cell_around_point per point)."""
        q = _c(np.asarray(points, dtype=np.float64).reshape(-1, 3), np.float64)
        out = np.empty(q.shape[0], dtype=np.int32)
        self._check(self._L.emfem_locate_points(self._h, q.shape[0], _p(q), _p(out)))
        return out

    def update_background_model(self):
        self._check(self._L.emfem_update_background_model(self._h))

    # -- linear system -----------------------------------------------------------
    def create_linear_system(self, n_slots):
        self._check(self._L.emfem_create_linear_system(self._h, int(n_slots)))

    def destroy_linear_system(self):
        self._check(self._L.emfem_destroy_linear_system(self._h))

    def assemble_matrix(self, fidx):
        self._check(self._L.emfem_assemble_matrix(self._h, int(fidx)))

    def get_matrix(self):
        inf = self.mesh_info()
        rp = np.empty(inf["n_edges"] + 1, dtype=np.int32)
        ci = np.empty(inf["nnz"], dtype=np.int32)
        C = np.empty(inf["nnz"])
        M = np.empty(inf["nnz"])
        self._check(self._L.emfem_get_matrix(self._h, _p(rp), _p(ci), _p(C), _p(M)))
        return rp, ci, C, M

    def assemble_rhs_mt(self, fidx, mode, slot):
        self._check(self._L.emfem_assemble_rhs_mt(self._h, int(fidx), int(mode), int(slot)))

    def assemble_rhs_csem(self, fidx, tidx, slot):
        self._check(self._L.emfem_assemble_rhs_csem(self._h, int(fidx), int(tidx), int(slot)))

    def assemble_rhs_dual(self, slot):
        self._check(self._L.emfem_assemble_rhs_dual(self._h, int(slot)))

    def set_rhs(self, slot, s):
        s = np.asarray(s)
        re = _c(s.real, np.float64)
        im = _c(s.imag, np.float64)
        self._check(self._L.emfem_set_rhs(self._h, int(slot), _p(re), _p(im)))

    def get_rhs(self, slot):
        n = self.mesh_info()["n_edges"]
        re = np.empty(n)
        im = np.empty(n)
        self._check(self._L.emfem_get_rhs(self._h, int(slot), _p(re), _p(im)))
        return re, im

    def get_solution(self, slot):
        n = self.mesh_info()["n_edges"]
        re = np.empty(n)
        im = np.empty(n)
        self._check(self._L.emfem_get_solution(self._h, int(slot), _p(re), _p(im)))
        return re + 1j * im

    def create_pc(self):
        self._check(self._L.emfem_create_pc(self._h))

    def prepare_pc(self, fidx):
        """Start the set-up for frequency fidx ahead of time (emfem_prepare_pc): the
        next assemble_matrix(fidx) and create_pc take the prepared matrix and host
        AMG set-up."""
        self._check(self._L.emfem_prepare_pc(self._h, int(fidx)))

    def destroy_pc(self):
        self._check(self._L.emfem_destroy_pc(self._h))

    def solve_linear_system(self, first, n, max_it, rtol, rtols=None, check=True, precision=None):
        """Returns (iterations, true relative residuals); raises on non-convergence
        when check is True (the reference never checks, src/em_la.cc:51-52).
        precision ('fp32' | 'fp64') overrides -inner_precision for this call."""
        its = np.zeros(n, dtype=np.int32)
        rr = np.zeros(n)
        rt = _c(rtols, np.float64) if rtols is not None else None
        keep = None
        if precision is not None:
            keep = self.get_option("inner_precision")
            self.process_options("-inner_precision %s" % precision)
        try:
            rc = self._L.emfem_solve_linear_system(self._h, int(first), int(n), int(max_it), float(rtol),
                                                   _p(rt) if rt is not None else None, _p(its), _p(rr))
        finally:
            if keep is not None:
                self.process_options("-inner_precision %s" % keep)
        if rc == ERR_NOT_CONVERGED and not check:
            return its, rr
        self._check(rc)
        return its, rr

    def matshell_mult_a(self, n_rhs, x_ptr, y_ptr):
        """y = A x on device pointers ([E][2 n_rhs] interleaved, fp64)."""
        self._check(self._L.emfem_matshell_mult_a(self._h, int(n_rhs), P(x_ptr), P(y_ptr)))

    def pc_apply_b(self, n_rhs, b_ptr, x_ptr):
        self._check(self._L.emfem_pc_apply_b(self._h, int(n_rhs), P(b_ptr), P(x_ptr)))

    def create_vecs(self, n_rhs, layout=LAYOUT_SPLIT):
        """Zeroed device vectors of the operator (emfem_create_vecs): (re, im) device
        pointers for LAYOUT_SPLIT, (x, None) for LAYOUT_INTERLEAVED; free each with
        device_free."""
        a, b = P(), P()
        self._check(self._L.emfem_create_vecs(self._h, int(n_rhs), int(layout), ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def device_free(self, ptr):
        self._check(self._L.emfem_device_free(self._h, P(ptr)))

    def copy_to_device(self, ptr, host):
        h = np.ascontiguousarray(host)
        self._check(self._L.emfem_copy_to_device(self._h, P(ptr), _p(h), ctypes.c_ulonglong(h.nbytes)))

    def copy_to_host(self, ptr, out):
        self._check(self._L.emfem_copy_to_host(self._h, _p(out), P(ptr), ctypes.c_ulonglong(out.nbytes)))
        return out

    def matshell_mult_a_split(self, n_rhs, xr, xi, yr, yi):
        """y = A x on the VecNest halves (split device arrays of n_rhs * n_edges)."""
        self._check(self._L.emfem_matshell_mult_a_split(self._h, int(n_rhs), P(xr), P(xi), P(yr), P(yi)))

    def matshell_mult_a_split_noalloc(self, n_rhs, xr, xi, yr, yi):
        """matshell_mult_a_split staged in the library's fixed device array (allocates
        nothing; n_edges * n_rhs <= 2097152)."""
        self._check(self._L.emfem_matshell_mult_a_split_noalloc(self._h, int(n_rhs), P(xr), P(xi), P(yr), P(yi)))

    def pc_apply_b_split(self, n_rhs, br, bi, xr, xi):
        self._check(self._L.emfem_pc_apply_b_split(self._h, int(n_rhs), P(br), P(bi), P(xr), P(xi)))

    # -- estimator and responses ----------------------------------------------------
    def estimate_error(self, fidx, slots):
        nt = self.mesh_info()["n_tets"]
        sl = _c(slots, np.int32)
        err = np.empty((sl.size, nt))
        self._check(self._L.emfem_estimate_error(self._h, int(fidx), sl.size, _p(sl), _p(err)))
        return err

    def interpolate_fields(self, fidx, slot, points):
        pts = _c(np.asarray(points).reshape(-1, 3), np.float64)
        E = np.empty((pts.shape[0], 3, 2))
        H = np.empty((pts.shape[0], 3, 2))
        self._check(self._L.emfem_interpolate_fields(self._h, int(fidx), int(slot), pts.shape[0], _p(pts), _p(E),
                                                     _p(H)))
        return E[..., 0] + 1j * E[..., 1], H[..., 0] + 1j * H[..., 1]

    def calculate_response_mt(self, fidx, slot_xy, slot_yx, obs):
        ob = _c(obs, np.int32)
        rsp = np.full((ob.shape[0], 2), np.nan)
        self._check(self._L.emfem_calculate_response_mt(self._h, int(fidx), int(slot_xy), int(slot_yx), ob.shape[0],
                                                        _p(ob), _p(rsp)))
        return rsp[:, 0] + 1j * rsp[:, 1]

    def calculate_response_csem(self, fidx, tidx, slot, obs):
        ob = _c(obs, np.int32)
        rsp = np.full((ob.shape[0], 2), np.nan)
        self._check(self._L.emfem_calculate_response_csem(self._h, int(fidx), int(tidx), int(slot), ob.shape[0],
                                                          _p(ob), _p(rsp)))
        return rsp[:, 0] + 1j * rsp[:, 1]

    # -- inputs, outputs, refinement, forward driver -----------------------------------
    def read_mesh(self, path):
        self._check(self._L.emfem_read_mesh(self._h, str(path).encode()))

    def read_emd(self, path):
        self._check(self._L.emfem_read_emd(self._h, str(path).encode()))

    def set_observations(self, obs, values=None):
        ob = _c(np.asarray(obs).reshape(-1, 4), np.int32)
        va = _c(np.asarray(values).reshape(-1, 4), np.float64) if values is not None else None
        self._check(self._L.emfem_set_observations(self._h, ob.shape[0], _p(ob), _p(va) if va is not None else None))
        self._nobs = ob.shape[0]

    def save_rsp(self, path):
        self._check(self._L.emfem_save_rsp(self._h, str(path).encode()))

    def save_mesh(self, path, tidx, error0, error1=None):
        """save_mesh (src/em_io.cc:289-314): legacy VTK with error indicators
        (MT: tidx < 0, error0 = XY, error1 = YX; CSEM: error0)."""
        e0 = _c(error0, np.float64)
        e1 = _c(error1, np.float64) if error1 is not None else None
        self._check(self._L.emfem_save_mesh(self._h, str(path).encode(), int(tidx), _p(e0),
                                            _p(e1) if e1 is not None else None))

    # -- learned indicator surrogate ----------------------------------------------
    def surrogate_init(self, seed=0):
        self._check(self._L.emfem_surrogate_init(self._h, int(seed)))

    def surrogate_params(self, set_params=None):
        """Returns the flat parameter vector, or installs set_params."""
        n = ctypes.c_int(0)
        self._check(self._L.emfem_surrogate_params(self._h, None, None, ctypes.byref(n)))
        if set_params is not None:
            sp = _c(set_params, np.float32)
            assert sp.size == n.value
            self._check(self._L.emfem_surrogate_params(self._h, None, _p(sp), ctypes.byref(n)))
            return sp
        out = np.empty(n.value, dtype=np.float32)
        self._check(self._L.emfem_surrogate_params(self._h, _p(out), None, ctypes.byref(n)))
        return out

    def surrogate_features(self, fidx, slot_xy, slot_yx):
        f = np.empty((self.mesh_info()["n_tets"], 16), dtype=np.float32)
        self._check(self._L.emfem_surrogate_features(self._h, int(fidx), int(slot_xy), int(slot_yx), _p(f)))
        return f

    def surrogate_train(self, feats, neighbors, indicator, epochs, lr=1e-3):
        f = _c(feats, np.float32)
        nb = _c(neighbors, np.int32)
        ind = _c(indicator, np.float64)
        loss = np.zeros(int(epochs))
        self._check(self._L.emfem_surrogate_train(self._h, f.shape[0], _p(f), _p(nb), _p(ind), int(epochs),
                                                  float(lr), _p(loss)))
        return loss

    def surrogate_predict(self, fidx, slot_xy, slot_yx):
        out = np.empty(self.mesh_info()["n_tets"])
        self._check(self._L.emfem_surrogate_predict(self._h, int(fidx), int(slot_xy), int(slot_yx), _p(out)))
        return out

    def surrogate_eval(self, feats, neighbors, target=None):
        """(y, loss, grad) of the raw network on standardized features."""
        f = _c(feats, np.float32)
        nb = _c(neighbors, np.int32)
        n = f.shape[0]
        y = np.empty(n, dtype=np.float32)
        loss = np.zeros(1)
        npar = ctypes.c_int(0)
        self._check(self._L.emfem_surrogate_params(self._h, None, None, ctypes.byref(npar)))
        grad = np.empty(npar.value - 34, dtype=np.float32)
        tg = _c(target, np.float32) if target is not None else None
        self._check(self._L.emfem_surrogate_eval(self._h, n, _p(f), _p(nb), _p(tg), _p(y),
                                                 _p(loss) if tg is not None else None,
                                                 _p(grad) if tg is not None else None))
        return (y, float(loss[0]), grad) if tg is not None else (y, None, None)

    def responses(self, n_obs=None):
        n = self._nobs if n_obs is None else n_obs
        r = np.empty((n, 2))
        self._check(self._L.emfem_get_responses(self._h, _p(r)))
        return r[:, 0] + 1j * r[:, 1]

    def forward(self):
        self._check(self._L.emfem_forward(self._h))

    def refine_mesh(self, flags):
        f = _c(flags, np.uint8)
        self._check(self._L.emfem_refine_mesh(self._h, _p(f)))

    def refine_uniform(self):
        self._check(self._L.emfem_refine_uniform(self._h))

    def get_mesh(self):
        inf = self.mesh_info()
        pts = np.empty((inf["n_vertices"], 3))
        tets = np.empty((inf["n_tets"], 4), dtype=np.int32)
        attr = np.empty(inf["n_tets"], dtype=np.int32)
        self._check(self._L.emfem_get_mesh(self._h, _p(pts), _p(tets), _p(attr)))
        return pts, tets, attr

    # -- device buffers -----------------------------------------------------------------
    def device_array(self, host):
        """A DeviceArray holding a copy of the float64 host array."""
        return DeviceArray(self, np.ascontiguousarray(host, dtype=np.float64))

    # -- measurement ------------------------------------------------------------------
    def stat(self, name):
        v = D()
        self._check(self._L.emfem_get_stat(self._h, name.encode(), ctypes.byref(v)))
        return v.value

    def kernel_time(self, which):
        ms = D()
        by = D()
        self._check(self._L.emfem_kernel_time(self._h, which.encode(), ctypes.byref(ms), ctypes.byref(by)))
        return ms.value, by.value


class DeviceArray:
    """A float64 device buffer owned through the context (emfem_device_alloc)."""

    def __init__(self, ctx, host):
        self.ctx = ctx
        self.shape = host.shape
        self.nbytes = host.nbytes
        p = P()
        ctx._check(ctx._L.emfem_device_alloc(ctx._h, ctypes.c_ulonglong(self.nbytes), ctypes.byref(p)))
        self.ptr = p.value
        ctx._check(ctx._L.emfem_copy_to_device(ctx._h, P(self.ptr), _p(host), ctypes.c_ulonglong(self.nbytes)))

    def numpy(self):
        out = np.empty(self.shape)
        self.ctx._check(self.ctx._L.emfem_copy_to_host(self.ctx._h, _p(out), P(self.ptr),
                                                       ctypes.c_ulonglong(self.nbytes)))
        return out

    def free(self):
        if self.ptr and getattr(self.ctx, "_h", None):
            self.ctx._L.emfem_device_free(self.ctx._h, P(self.ptr))
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def refine_fixed_fraction(error, fraction, flag=None):
    e = _c(error, np.float64)
    f = np.zeros(e.size, dtype=np.uint8) if flag is None else flag
    rc = lib().emfem_refine_fixed_fraction(e.size, _p(e), float(fraction), _p(f))
    if rc != OK:
        raise EmfemError(rc, "refine_fixed_fraction")
    return f


def refine_fixed_number(error, fraction, flag=None):
    e = _c(error, np.float64)
    f = np.zeros(e.size, dtype=np.uint8) if flag is None else flag
    rc = lib().emfem_refine_fixed_number(e.size, _p(e), float(fraction), _p(f))
    if rc != OK:
        raise EmfemError(rc, "refine_fixed_number")
    return f


def solve_batch(contexts, first, n, max_it, rtol, check=True):
    """Multi-model batch solve (emfem_solve_batch): slots [first, first+n) of every
    context as one block-diagonal system with per-model Krylov processes.  Returns
    (iterations [k, n], true relative residuals [k, n], acceptance times in ms [k, n]).
    Options are taken from contexts[0]."""
    k = len(contexts)
    hs = (P * k)(*[c._h for c in contexts])
    its = np.zeros((k, n), dtype=np.int32)
    rr = np.zeros((k, n))
    ms = np.zeros((k, n))
    rc = lib().emfem_solve_batch(hs, k, int(first), int(n), int(max_it), float(rtol), _p(its), _p(rr), _p(ms))
    if rc != OK and (check or rc != ERR_NOT_CONVERGED):
        contexts[0]._check(rc)
    return its, rr, ms
