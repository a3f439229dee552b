"""ctypes wrapper of the oracle library (oracle/_build/libpm_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, never by the product package.  Consumes the flat problem
description produced by ``pmsim_amd.problem.describe``.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libpm_oracle.so")
_lib = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_lp = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.orc_new.restype = C.c_void_p
        L.orc_new.argtypes = [C.c_int, C.c_int, _dp, C.c_int, _ip, _ip, _ip, _ip]
        L.orc_free.argtypes = [C.c_void_p]
        L.orc_last_error.restype = C.c_char_p
        L.orc_last_error.argtypes = [C.c_void_p]
        L.orc_set_param.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int]
        L.orc_set_ideal.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double]
        L.orc_set_config.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int, C.c_double, C.c_double]
        L.orc_set_essential.argtypes = [C.c_void_p, C.c_int, _ip, _ip, _dp, C.c_int]
        L.orc_set_point_loads.argtypes = [C.c_void_p, C.c_int, _ip, _ip, _dp, C.c_int]
        L.orc_add_facet_qn.argtypes = [C.c_void_p, C.c_int, C.c_int, _ip, C.c_int, _dp, C.c_int]
        L.orc_add_facet_load.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, _ip, C.c_int, _dp, C.c_int]
        L.orc_f_ext.argtypes = [C.c_void_p, C.c_double, _dp]
        L.orc_setup.argtypes = [C.c_void_p]
        L.orc_sizes.argtypes = [C.c_void_p, _ip]
        L.orc_solve.argtypes = [C.c_void_p, _dp, _dp, C.c_int, C.POINTER(C.c_int), _ip, C.c_int, _dp]
        L.orc_set_arc_length.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double]
        L.orc_init_state.argtypes = [C.c_void_p]
        L.orc_set_ip_state.argtypes = [C.c_void_p, _dp]
        L.orc_get_ip_state.argtypes = [C.c_void_p, _dp]
        L.orc_update_secondary.argtypes = [C.c_void_p, _dp]
        L.orc_assemble_f_int.argtypes = [C.c_void_p, _dp]
        L.orc_element_K.argtypes = [C.c_void_p, C.c_int, _dp]
        L.orc_element_f.argtypes = [C.c_void_p, C.c_int, _dp]
        L.orc_assemble_kk_csr.argtypes = [C.c_void_p, _lp, _ip, C.c_void_p, C.POINTER(C.c_int64)]
        L.orc_time_element_pass.restype = C.c_double
        L.orc_time_element_pass.argtypes = [C.c_void_p, _dp, C.c_int, C.c_int]
        L.orc_shape.argtypes = [C.c_int, _dp, _dp, _dp]
        L.orc_gauss.argtypes = [C.c_int, C.c_int, _dp]
        L.orc_material_update.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, _dp, C.c_void_p, C.c_void_p]
        _lib = L
    return _lib


class OracleError(RuntimeError):
    pass


class Oracle:
    """One problem held by the oracle library."""

    def __init__(self, d: dict):
        L = lib()
        self.d = d
        cp = np.ascontiguousarray(d["cell_ptr"], dtype=np.int32)
        self.h = L.orc_new(d["ndim"], d["coords"].shape[0], np.ascontiguousarray(d["coords"], dtype=np.float64).ravel(),
                           d["kinds"].shape[0], np.ascontiguousarray(d["kinds"], dtype=np.int32),
                           np.ascontiguousarray(d["attrs"], dtype=np.int32), cp,
                           np.ascontiguousarray(d["cell_pts"], dtype=np.int32))
        for a, (model, E, nu, H, z0, ng) in d["params"].items():
            L.orc_set_param(self.h, int(a), int(model), E, nu, H, z0, int(ng))
        L.orc_set_ideal(self.h, int(d["axisym"]), int(d["plane_stress"]), float(d["thickness"]))
        L.orc_set_config(self.h, int(d["lmm"]), int(d["linear_problem"]), d["t_ini"], d["t_fin"], d["dt"],
                         int(d["n_max_iterations"]), d["tol_rr_abs"], d["tol_mdu_rel"])
        if d.get("arc_length_method"):
            L.orc_set_arc_length(self.h, 1, float(d["arc_length_psi"]), float(d["first_trial_loading_factor"]))
        tab = np.ascontiguousarray(d["ess_table"], dtype=np.float64)
        n = tab.shape[0]
        L.orc_set_essential(self.h, n, np.ascontiguousarray(d["ess_points"], dtype=np.int32),
                            np.ascontiguousarray(d["ess_dofs"], dtype=np.int32), tab.ravel() if n else np.zeros(1),
                            tab.shape[1] if n else 1)
        plt = np.ascontiguousarray(d["pl_table"], dtype=np.float64)
        if plt.shape[0]:
            L.orc_set_point_loads(self.h, plt.shape[0], d["pl_points"], d["pl_dofs"], plt.ravel(), plt.shape[1])
        for facet in d["facets"]:
            fk, pts, cell, v = facet[:4]
            nbc = facet[4] if len(facet) > 4 else 0
            p = np.asarray(pts, dtype=np.int32)
            v = np.ascontiguousarray(v, dtype=np.float64)
            L.orc_add_facet_load(self.h, int(nbc), int(fk), len(p), p, int(cell), v, len(v))
        sz = np.zeros(4, dtype=np.int32)
        self._check(L.orc_sizes(self.h, sz))
        self.ndof, self.neq_total, self.nip, self.ns = (int(v) for v in sz)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_free(self.h)
            self.h = None

    def _check(self, rc):
        if rc != 0:
            raise OracleError(lib().orc_last_error(self.h).decode())

    def solve(self, max_out=None):
        """Runs the Newton driver; returns (u_hist [nout×neq_total], ip_hist [nout×nip×(ns+4)], iters).
        The loading factor of each output is kept in ``last_ell``; with the arc-length method a step that
        did not converge reports minus its iteration count."""
        L = lib()
        nst = len(np.arange(self.d["t_ini"], self.d["t_fin"] + 1e-9 * self.d["dt"], self.d["dt"])) + 1
        max_out = max_out or nst + 1
        u = np.zeros(max_out * self.neq_total)
        ip = np.zeros(max_out * self.nip * (self.ns + 4))
        nout = C.c_int(0)
        iters = np.zeros(max(nst, 1) + 1, dtype=np.int32)
        ell = np.zeros(max_out)
        rc = L.orc_solve(self.h, u, ip, max_out, C.byref(nout), iters, len(iters), ell)
        n = min(nout.value, max_out)
        self.last_nout = nout.value
        self.last_ell = ell[:n]
        self.last_u = u.reshape(max_out, self.neq_total)[:n]
        self.last_ip = ip.reshape(max_out, self.nip, self.ns + 4)[:n]
        self.last_iters = iters
        self._check(rc)
        return self.last_u, self.last_ip, iters[: max(n - 1, 0)]

    # element-level access
    def init_state(self):
        self._check(lib().orc_init_state(self.h))

    def set_ip_state(self, v):
        lib().orc_set_ip_state(self.h, np.ascontiguousarray(v, dtype=np.float64).ravel())

    def get_ip_state(self):
        v = np.zeros(self.nip * (self.ns + 4))
        lib().orc_get_ip_state(self.h, v)
        return v.reshape(self.nip, self.ns + 4)

    def update_secondary(self, ddu):
        self._check(lib().orc_update_secondary(self.h, np.ascontiguousarray(ddu, dtype=np.float64)))

    def f_int(self):
        f = np.zeros(self.ndof)
        lib().orc_assemble_f_int(self.h, f)
        return f

    def f_ext(self, t):
        """F_ext at time t (distributed and point loads; prescribed equations skipped under RSM)."""
        f = np.zeros(self.ndof)
        self._check(lib().orc_f_ext(self.h, float(t), f))
        return f

    def element_f(self, cell, ne):
        f = np.zeros(ne)
        self._check(lib().orc_element_f(self.h, int(cell), f))
        return f

    def element_K(self, cell, ne):
        K = np.zeros(ne * ne)
        self._check(lib().orc_element_K(self.h, int(cell), K))
        return K.reshape(ne, ne)

    def kk_csr(self):
        L = lib()
        nmax = int(np.sum(np.diff(self.d["cell_ptr"]) ** 2)) * self.d["ndim"] ** 2
        rowptr = np.zeros(self.ndof + 1, dtype=np.int64)
        cols = np.zeros(nmax, dtype=np.int32)
        nnz = C.c_int64(0)
        vals = np.zeros(nmax)
        self._check(L.orc_assemble_kk_csr(self.h, rowptr, cols, vals.ctypes.data_as(C.c_void_p), C.byref(nnz)))
        k = nnz.value
        return rowptr, cols[:k].copy(), vals[:k].copy()

    def time_element_pass(self, ddu, reps=1, nthreads=1):
        return lib().orc_time_element_pass(self.h, np.ascontiguousarray(ddu, dtype=np.float64), int(reps), int(nthreads))


def shape(kind, x):
    x = np.asarray(list(x) + [0.0] * (3 - len(x)), dtype=np.float64)
    N = np.zeros(32)
    dN = np.zeros(96)
    n = lib().orc_shape(int(kind), x, N, dN)
    return N[:n].copy(), dN


def gauss(kind, ngauss=0):
    xw = np.zeros(64 * 4)
    n = lib().orc_gauss(int(kind), int(ngauss), xw)
    return xw.reshape(-1, 4)[:n].copy()


def material_update(ndim, model, E, nu, H, z0, state, deps=None, want_D=False):
    ns = 4 if ndim == 2 else 6
    st = np.ascontiguousarray(state, dtype=np.float64).copy()
    D = np.zeros(ns * ns) if want_D else None
    de = np.ascontiguousarray(deps, dtype=np.float64) if deps is not None else None
    rc = lib().orc_material_update(ndim, model, E, nu, H, z0, st,
                                   de.ctypes.data_as(C.c_void_p) if de is not None else None,
                                   D.ctypes.data_as(C.c_void_p) if D is not None else None)
    return rc, st, (D.reshape(ns, ns) if D is not None else None)
