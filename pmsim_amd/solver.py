"""ctypes binding of libpmsim_amd.so and the pmsim-style solver facade.

``SolverImplicit(mesh, base, config, essential, natural).solve(state)`` mirrors pmsim's
``SolverImplicit::new`` / ``solve`` (src/fem/solver_implicit.rs:27-178); ``FemState`` holds the
primary variables and the per-IP secondary values (src/fem/fem_state.rs:15-96).  The compute path
is the HIP library: if it is not built, importing this module raises (there is no CPU fallback).
"""
from __future__ import annotations

import ctypes as C
import os
from typing import List, Optional

import numpy as np

from .problem import Config, Essential, FemBase, Natural, describe
from . import mesh as M
from .mesh import Mesh

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpmsim_amd.so")

EXPORTED = [
    "pm_last_error", "pm_global_last_error", "pm_default_config", "pm_fem_create", "pm_fem_destroy", "pm_set_param_solid",
    "pm_set_idealization", "pm_set_config", "pm_set_essential", "pm_set_point_loads", "pm_add_facet_qn", "pm_initialize",
    "pm_initialize_internal_values", "pm_reset_algorithmic_variables", "pm_backup_secondary_values",
    "pm_restore_secondary_values", "pm_update_secondary_values", "pm_assemble_f_int", "pm_assemble_kk", "pm_element_pass",
    "pm_solve", "pm_get_stats", "pm_sizes", "pm_set_ddu", "pm_get_u", "pm_get_f_int", "pm_get_ip_state", "pm_set_ip_state",
    "pm_get_csr", "pm_history_count", "pm_get_history", "pm_time_element_pass", "pm_linear_solve",
    "pm_estimate_device_bytes", "pm_release_device", "pm_serve", "pm_gauss_rule", "pm_extrapolation_matrix",
    "pm_shape_values", "pm_get_history_loading_factor", "pm_add_facet_load", "pm_assemble_f_ext", "pm_get_f_ext",
    "pm_calculate_residuals_vector", "pm_get_residual", "pm_apply_kk", "pm_get_element_kk", "pm_get_element_f_int",
    "pm_last_error_location", "pm_linear_solve_noalloc", "pm_server_create", "pm_server_submit", "pm_server_poll",
    "pm_server_wait", "pm_server_destroy", "pm_server_stats", "pm_set_model_settings",
]


class PmConfig(C.Structure):
    _fields_ = [
        ("lagrange_mult_method", C.c_int32), ("linear_problem", C.c_int32), ("constant_tangent", C.c_int32),
        ("n_max_time_steps", C.c_int32), ("n_max_iterations", C.c_int32), ("allowed_step_n_failure", C.c_int32),
        ("t_ini", C.c_double), ("t_fin", C.c_double), ("dt", C.c_double), ("dt_out", C.c_double),
        ("tol_rr_abs", C.c_double), ("tol_mdu_rel", C.c_double), ("symmetry_check_tolerance", C.c_double),
        ("lin_sol", C.c_int32), ("tangent_precision", C.c_int32), ("pcg_rtol", C.c_double), ("pcg_max_iter", C.c_int32),
        ("keep_history", C.c_int32), ("tol_rr_scaled", C.c_double), ("tangent_kernel", C.c_int32),
        ("use_graphs", C.c_int32), ("pcg_small_max", C.c_int32), ("verbose", C.c_int32),
        ("arc_length_method", C.c_int32), ("arc_length_psi", C.c_double), ("first_trial_loading_factor", C.c_double),
    ]


class PmRequestStats(C.Structure):
    _fields_ = [
        ("t_arrival", C.c_double), ("t_admit", C.c_double), ("t_done", C.c_double), ("status", C.c_int32),
        ("n_steps", C.c_int32), ("n_iterations", C.c_int32), ("n_pcg_iterations", C.c_int64), ("device_bytes", C.c_int64),
    ]


class PmStats(C.Structure):
    _fields_ = [
        ("n_steps", C.c_int32), ("n_iterations", C.c_int32), ("n_pcg_iterations", C.c_int64),
        ("n_host_fallbacks", C.c_int32), ("ms_update", C.c_double), ("ms_residual", C.c_double),
        ("ms_tangent", C.c_double), ("ms_scatter", C.c_double), ("ms_linear", C.c_double), ("ms_total", C.c_double),
        ("n_ip", C.c_int64), ("n_block_nnz", C.c_int64), ("device_bytes", C.c_int64),
    ]


_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_lp = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_lib = None


def lib():
    """Loads the HIP library; raises if it has not been built (no silent fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built: run `python -m pmsim_amd.build`")
    L = C.CDLL(LIB_PATH)
    L.pm_last_error.restype = C.c_char_p
    L.pm_last_error.argtypes = [C.c_void_p]
    L.pm_global_last_error.restype = C.c_char_p
    L.pm_default_config.argtypes = [C.POINTER(PmConfig)]
    L.pm_fem_create.argtypes = [C.POINTER(C.c_void_p), C.c_int32, C.c_int64, _dp, C.c_int64, _ip, _ip, _lp, _ip]
    L.pm_fem_destroy.argtypes = [C.c_void_p]
    L.pm_set_param_solid.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int32]
    L.pm_set_model_settings.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_double]
    L.pm_set_idealization.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_double]
    L.pm_set_config.argtypes = [C.c_void_p, C.POINTER(PmConfig)]
    L.pm_set_essential.argtypes = [C.c_void_p, C.c_int64, _ip, _ip, _dp, C.c_int32]
    L.pm_set_point_loads.argtypes = [C.c_void_p, C.c_int64, _ip, _ip, _dp, C.c_int32]
    L.pm_add_facet_qn.argtypes = [C.c_void_p, C.c_int32, C.c_int32, _ip, C.c_int64, _dp, C.c_int32]
    L.pm_add_facet_load.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, _ip, C.c_int64, _dp, C.c_int32]
    L.pm_assemble_f_ext.argtypes = [C.c_void_p, C.c_double]
    L.pm_get_f_ext.argtypes = [C.c_void_p, _dp]
    L.pm_calculate_residuals_vector.argtypes = [C.c_void_p, C.c_double]
    L.pm_get_residual.argtypes = [C.c_void_p, _dp]
    L.pm_apply_kk.argtypes = [C.c_void_p, _dp, _dp]
    L.pm_get_element_kk.argtypes = [C.c_void_p, C.c_int64, _dp]
    L.pm_get_element_f_int.argtypes = [C.c_void_p, C.c_int64, _dp]
    L.pm_last_error_location.argtypes = [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int32)]
    L.pm_initialize.argtypes = [C.c_void_p, C.c_int32]
    for name in ["pm_initialize_internal_values", "pm_reset_algorithmic_variables", "pm_backup_secondary_values",
                 "pm_restore_secondary_values", "pm_update_secondary_values", "pm_assemble_f_int", "pm_assemble_kk",
                 "pm_element_pass"]:
        getattr(L, name).argtypes = [C.c_void_p]
    L.pm_solve.argtypes = [C.c_void_p, C.POINTER(C.c_int32)]
    L.pm_get_stats.argtypes = [C.c_void_p, C.POINTER(PmStats)]
    L.pm_sizes.argtypes = [C.c_void_p, _lp]
    L.pm_set_ddu.argtypes = [C.c_void_p, _dp]
    L.pm_get_u.argtypes = [C.c_void_p, _dp, C.c_void_p]
    L.pm_get_f_int.argtypes = [C.c_void_p, _dp]
    L.pm_get_ip_state.argtypes = [C.c_void_p, _dp]
    L.pm_set_ip_state.argtypes = [C.c_void_p, _dp]
    L.pm_get_csr.argtypes = [C.c_void_p, _lp, _ip, _dp]
    L.pm_history_count.argtypes = [C.c_void_p, C.POINTER(C.c_int32)]
    L.pm_get_history.argtypes = [C.c_void_p, C.c_int32, _dp, C.c_void_p, _dp, C.POINTER(C.c_double)]
    L.pm_get_history_loading_factor.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.c_double)]
    L.pm_time_element_pass.argtypes = [C.c_void_p, C.c_int32, _dp]
    L.pm_gauss_rule.argtypes = [C.c_int32, C.c_int32, C.POINTER(C.c_int32), C.c_void_p, C.c_void_p]
    L.pm_extrapolation_matrix.argtypes = [C.c_int32, C.c_int32, _dp]
    L.pm_shape_values.argtypes = [C.c_int32, C.c_int32, _dp, _dp]
    L.pm_linear_solve.argtypes = [C.c_void_p, _dp, _dp, C.POINTER(C.c_int32)]
    L.pm_estimate_device_bytes.argtypes = [C.c_void_p, C.POINTER(C.c_int64)]
    L.pm_release_device.argtypes = [C.c_void_p]
    L.pm_serve.argtypes = [C.c_int32, C.POINTER(C.c_void_p), _dp, C.c_int32, C.c_int64, C.c_int32, C.c_void_p]
    L.pm_server_create.argtypes = [C.POINTER(C.c_void_p), C.c_int32, C.c_int64, C.c_int32]
    L.pm_server_submit.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(C.c_int64)]
    L.pm_server_poll.argtypes = [C.c_void_p, C.c_int64, C.c_void_p]
    L.pm_server_wait.argtypes = [C.c_void_p, C.c_int64, C.c_void_p]
    L.pm_server_destroy.argtypes = [C.c_void_p]
    L.pm_server_stats.argtypes = [C.c_void_p, _lp]
    _lib = L
    return L


class PmError(RuntimeError):
    pass


PRECISION = {"f64": 0, "f32": 1, "bf16": 2}
TANGENT_KERNEL = {"default": 0, "vector": 1, "mfma": 2, "tune": 3, "vector-wg": 4, "mfma-wg": 5}


def make_config(d: dict, cfg: Optional[Config] = None, keep_history=True) -> PmConfig:
    c = PmConfig()
    lib().pm_default_config(C.byref(c))
    c.lagrange_mult_method = int(d["lmm"])
    c.linear_problem = int(d["linear_problem"])
    c.n_max_iterations = int(d["n_max_iterations"])
    c.t_ini, c.t_fin, c.dt, c.dt_out = d["t_ini"], d["t_fin"], d["dt"], d["dt"]
    c.tol_rr_abs, c.tol_mdu_rel = d["tol_rr_abs"], d["tol_mdu_rel"]
    c.keep_history = 1 if keep_history else 0
    c.arc_length_method = 1 if d.get("arc_length_method") else 0
    c.arc_length_psi = float(d.get("arc_length_psi", 1.0))
    c.first_trial_loading_factor = float(d.get("first_trial_loading_factor", 1.0))
    if cfg is not None:
        c.constant_tangent = int(cfg.constant_tangent)
        c.n_max_time_steps = int(cfg.n_max_time_steps)
        c.allowed_step_n_failure = int(cfg.allowed_step_n_failure)
        c.symmetry_check_tolerance = -1.0 if cfg.symmetry_check_tolerance is None else float(cfg.symmetry_check_tolerance)
        c.lin_sol = 0 if cfg.lin_sol == "pcg" else 1
        c.tangent_precision = PRECISION[cfg.tangent_precision]
        c.pcg_rtol = cfg.pcg_tol
        c.pcg_max_iter = cfg.pcg_max_iter
        c.tol_rr_scaled = cfg.scaled_rr
        c.tangent_kernel = TANGENT_KERNEL[cfg.tangent_kernel]
        c.use_graphs = 1 if cfg.use_graphs else 0
        c.pcg_small_max = int(cfg.pcg_small_max)
        c.verbose = (1 if cfg.verbose_timesteps else 0) | (2 if cfg.verbose_iterations else 0)
    return c


class Fem:
    """Owns one pm_fem handle built from a flat problem description (``problem.describe``)."""

    def __init__(self, d: dict, cfg: Optional[Config] = None, device: int = 0, keep_history=True, pm_config=None,
                 initialize=True):
        L = lib()
        self.d = d
        h = C.c_void_p()
        X = np.ascontiguousarray(d["coords"], dtype=np.float64).ravel()
        rc = L.pm_fem_create(C.byref(h), int(d["ndim"]), d["coords"].shape[0], X, d["kinds"].shape[0],
                             np.ascontiguousarray(d["kinds"], dtype=np.int32), np.ascontiguousarray(d["attrs"], dtype=np.int32),
                             np.ascontiguousarray(d["cell_ptr"], dtype=np.int64), np.ascontiguousarray(d["cell_pts"], dtype=np.int32))
        if rc != 0:
            raise PmError(L.pm_global_last_error().decode())
        self.h = h
        for a, (model, E, nu, H, z0, ng) in d["params"].items():
            self._check(L.pm_set_param_solid(h, int(a), int(model), E, nu, H, z0, int(ng)))
        for a, (gp, nn_max, beta) in d.get("model_settings", {}).items():
            self._check(L.pm_set_model_settings(h, int(a), int(gp), int(nn_max), float(beta)))
        self._check(L.pm_set_idealization(h, int(d["axisym"]), int(d["plane_stress"]), float(d["thickness"])))
        self.cfg = pm_config if pm_config is not None else make_config(d, cfg, keep_history)
        self._check(L.pm_set_config(h, C.byref(self.cfg)))
        tab = np.ascontiguousarray(d["ess_table"], dtype=np.float64)
        n = tab.shape[0]
        if n:
            self._check(L.pm_set_essential(h, n, np.ascontiguousarray(d["ess_points"], dtype=np.int32),
                                           np.ascontiguousarray(d["ess_dofs"], dtype=np.int32), tab.ravel(), tab.shape[1]))
        plt = np.ascontiguousarray(d["pl_table"], dtype=np.float64)
        if plt.shape[0]:
            self._check(L.pm_set_point_loads(h, plt.shape[0], np.ascontiguousarray(d["pl_points"], dtype=np.int32),
                                             np.ascontiguousarray(d["pl_dofs"], dtype=np.int32), plt.ravel(), plt.shape[1]))
        for facet in d["facets"]:
            fk, pts, cell, v = facet[:4]
            nbc = facet[4] if len(facet) > 4 else 0
            p = np.asarray(pts, dtype=np.int32)
            v = np.ascontiguousarray(v, dtype=np.float64)
            self._check(L.pm_add_facet_load(h, int(nbc), int(fk), len(p), p, int(cell), v, len(v)))
        self.device = device
        if initialize:
            self.initialize()

    def initialize(self):
        """Uploads the problem to the device (pm_initialize); done by the constructor unless it was
        built with initialize=False (serving initializes at admission)."""
        self._check(lib().pm_initialize(self.h, self.device))
        self.refresh_sizes()

    def refresh_sizes(self):
        sz = np.zeros(6, dtype=np.int64)
        self._check(lib().pm_sizes(self.h, sz))
        self.ndof, self.nprescribed, self.nip, self.ns, self.nnzb, self.ndim = (int(v) for v in sz)

    def estimate_device_bytes(self):
        b = C.c_int64(0)
        self._check(lib().pm_estimate_device_bytes(self.h, C.byref(b)))
        return b.value

    def release_device(self):
        """Keep u, lambda and the IP state on the host and free the handle's device memory."""
        self._check(lib().pm_release_device(self.h))

    def close(self):
        if getattr(self, "h", None):
            lib().pm_fem_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != 0:
            raise PmError(lib().pm_last_error(self.h).decode())

    # ElementTrait batched over device-resident state
    def initialize_internal_values(self): self._check(lib().pm_initialize_internal_values(self.h))
    def reset_algorithmic_variables(self): self._check(lib().pm_reset_algorithmic_variables(self.h))
    def backup_secondary_values(self): self._check(lib().pm_backup_secondary_values(self.h))
    def restore_secondary_values(self): self._check(lib().pm_restore_secondary_values(self.h))
    def update_secondary_values(self): self._check(lib().pm_update_secondary_values(self.h))
    def assemble_f_int(self): self._check(lib().pm_assemble_f_int(self.h))
    def assemble_kk(self): self._check(lib().pm_assemble_kk(self.h))
    def element_pass(self): self._check(lib().pm_element_pass(self.h))

    def set_ddu(self, ddu):
        self._check(lib().pm_set_ddu(self.h, np.ascontiguousarray(ddu, dtype=np.float64)))

    def get_u(self):
        u = np.zeros(self.ndof)
        lam = np.zeros(max(self.nprescribed, 1))
        self._check(lib().pm_get_u(self.h, u, lam.ctypes.data_as(C.c_void_p)))
        return u, lam[: self.nprescribed]

    def get_f_int(self):
        f = np.zeros(self.ndof)
        self._check(lib().pm_get_f_int(self.h, f))
        return f

    def assemble_f_ext(self, t):
        """F_ext at time t (SolverData::assemble_ff_int_and_ff_ext, solver_data.rs:85-108)."""
        self._check(lib().pm_assemble_f_ext(self.h, float(t)))

    def get_f_ext(self):
        f = np.zeros(self.ndof)
        self._check(lib().pm_get_f_ext(self.h, f))
        return f

    def calculate_residuals_vector(self, loading_factor=1.0):
        """R = F_int - loading_factor F_ext (solver_data.rs:111-120)."""
        self._check(lib().pm_calculate_residuals_vector(self.h, float(loading_factor)))

    def get_residual(self):
        r = np.zeros(self.ndof)
        self._check(lib().pm_get_residual(self.h, r))
        return r

    def apply_kk(self, x):
        """y = K x with the last assembled tangent."""
        y = np.zeros(self.ndof)
        self._check(lib().pm_apply_kk(self.h, np.ascontiguousarray(x, dtype=np.float64), y))
        return y

    def element_kk(self, cell):
        """Local tangent of a cell (ElementTrait::calc_jacobian), local DOF order ndim*m + d."""
        n = int(self.d["cell_ptr"][cell + 1] - self.d["cell_ptr"][cell]) * self.ndim
        k = np.zeros(n * n)
        self._check(lib().pm_get_element_kk(self.h, int(cell), k))
        return k.reshape(n, n)

    def element_f_int(self, cell):
        """Local internal force of a cell (ElementTrait::calc_f_int)."""
        n = int(self.d["cell_ptr"][cell + 1] - self.d["cell_ptr"][cell]) * self.ndim
        f = np.zeros(n)
        self._check(lib().pm_get_element_f_int(self.h, int(cell), f))
        return f

    def last_error_location(self):
        """(cell, gauss point) of the first IP that raised the last device-side error, or (-1, -1)."""
        c, g = C.c_int64(-1), C.c_int32(-1)
        lib().pm_last_error_location(self.h, C.byref(c), C.byref(g))
        return c.value, g.value

    def get_ip_state(self):
        v = np.zeros(self.nip * (self.ns + 4))
        self._check(lib().pm_get_ip_state(self.h, v))
        return v.reshape(self.nip, self.ns + 4)

    def set_ip_state(self, v):
        self._check(lib().pm_set_ip_state(self.h, np.ascontiguousarray(v, dtype=np.float64).ravel()))

    def get_csr(self):
        nnz = self.nnzb * self.ndim * self.ndim
        rowptr = np.zeros(self.ndof + 1, dtype=np.int64)
        cols = np.zeros(nnz, dtype=np.int32)
        vals = np.zeros(nnz)
        self._check(lib().pm_get_csr(self.h, rowptr, cols, vals))
        return rowptr, cols, vals

    def solve(self):
        n = C.c_int32(0)
        self._check(lib().pm_solve(self.h, C.byref(n)))
        return n.value

    def stats(self):
        s = PmStats()
        lib().pm_get_stats(self.h, C.byref(s))
        return {k: getattr(s, k) for (k, _) in PmStats._fields_}

    def history(self):
        L = lib()
        n = C.c_int32(0)
        L.pm_history_count(self.h, C.byref(n))
        out = []
        for i in range(n.value):
            u = np.zeros(self.ndof)
            lam = np.zeros(max(self.nprescribed, 1))
            ip = np.zeros(self.nip * (self.ns + 4))
            t = C.c_double(0)
            self._check(L.pm_get_history(self.h, i, u, lam.ctypes.data_as(C.c_void_p), ip, C.byref(t)))
            out.append((t.value, u, lam[: self.nprescribed], ip.reshape(self.nip, self.ns + 4)))
        return out

    def loading_factors(self):
        """Loading factor of every output step of the last solve (FemState::ell)."""
        L = lib()
        n = C.c_int32(0)
        L.pm_history_count(self.h, C.byref(n))
        out = np.zeros(n.value)
        for i in range(n.value):
            v = C.c_double(0)
            self._check(L.pm_get_history_loading_factor(self.h, i, C.byref(v)))
            out[i] = v.value
        return out

    def time_element_pass(self, reps=10):
        ms = np.zeros(6)
        self._check(lib().pm_time_element_pass(self.h, int(reps), ms))
        return dict(update=ms[0], residual=ms[1], tangent=ms[2], scatter=ms[3], total=ms[4], pass_=ms[5])

    def linear_solve(self, b):
        x = np.zeros(self.ndof)
        it = C.c_int32(0)
        self._check(lib().pm_linear_solve(self.h, np.ascontiguousarray(b, dtype=np.float64), x, C.byref(it)))
        return x, it.value


class FemState:
    """Primary variables and per-IP secondary values of the output steps (pmsim FemState)."""

    def __init__(self, mesh: Mesh, base: FemBase, essential: Essential, config: Config):
        self.t = 0.0
        self.ell = 0.0  # loading factor
        self.ell_history = np.zeros(0)
        self.u: Optional[np.ndarray] = None
        self.gauss: Optional[np.ndarray] = None  # nip × (nsig + 4): sigma (Mandel), z, lambda, elastic, epbar
        self.history: List[tuple] = []


class SolverImplicit:
    """pmsim's SolverImplicit over the HIP library."""

    def __init__(self, mesh: Mesh, base: FemBase, config: Config, essential: Essential, natural: Natural,
                 device: int = 0, keep_history: bool = True):
        if not config.lagrange_mult_method and essential.has_non_zero():
            # checked again by the library at solve time; raised here with pmsim's text for early failure
            pass
        self.mesh = mesh
        self.desc = describe(mesh, base, config, essential, natural)
        self.fem = Fem(self.desc, config, device=device, keep_history=keep_history)

    def cell_layout(self):
        """IPs and stress-strain model of every cell (cell order), as pmsim's state files hold them."""
        d = self.desc
        nip = [int(d["params"][int(a)][5]) or M.DEFAULT_NGAUSS[int(k)] for k, a in zip(d["kinds"], d["attrs"])]
        model = [int(d["params"][int(a)][0]) for a in d["attrs"]]
        return nip, model

    def write_files(self, file_io):
        """Mesh, base, one state file per output step and the summary, in pmsim's formats."""
        from .file_io import state_dict
        file_io.write_setup(self.mesh, self.desc)
        nip, model = self.cell_layout()
        prev = None
        for (t, u, lam, ip), ell in zip(self.fem.history(), self.fem.loading_factors()):
            ddu = u - prev if prev is not None else np.zeros_like(u)
            file_io.write_state(state_dict(self.mesh.ndim, t, float(self.desc["dt"]), u, ddu, ip, nip, model, ell=ell))
            prev = u
        file_io.write_summary()

    def solve(self, state: Optional[FemState] = None, file_io=None) -> int:
        n = self.fem.solve()
        if file_io is not None:
            self.write_files(file_io)
        if state is not None:
            state.history = self.fem.history()
            state.ell_history = self.fem.loading_factors()
            u, lam = self.fem.get_u()
            state.u = u
            state.gauss = self.fem.get_ip_state()
            if state.history:
                state.t = state.history[-1][0]
                state.ell = float(state.ell_history[-1])
        return n


def serve(fems, arrivals, device=0, hbm_budget_bytes=200 << 30, max_active=16):
    """Continuous batching of independent solves (pm_serve).  `fems` are handles built with
    initialize=False; request i arrives `arrivals[i]` seconds after the call.  Returns per-request
    dicts (arrival / admit / done times, status, iteration counts, footprint estimate)."""
    n = len(fems)
    arr = np.ascontiguousarray(arrivals, dtype=np.float64)
    hs = (C.c_void_p * n)(*[f.h for f in fems])
    out = (PmRequestStats * n)()
    if lib().pm_serve(n, hs, arr, int(device), int(hbm_budget_bytes), int(max_active), C.cast(out, C.c_void_p)) != 0:
        raise PmError(lib().pm_global_last_error().decode())
    res = []
    for f, o in zip(fems, out):
        r = {k: getattr(o, k) for (k, _) in PmRequestStats._fields_}
        r["error"] = lib().pm_last_error(f.h).decode() if o.status != 0 else ""
        if o.status == 0:
            f.refresh_sizes()
        res.append(r)
    return res


class Server:
    """Open-submission serving loop (pm_server_*): submit configured, not yet initialized handles while others
    run; poll or wait for each.  Element passes and PCG iterations of the running requests are batched."""

    def __init__(self, device=0, hbm_budget_bytes=200 << 30, max_active=16):
        self.h = C.c_void_p()
        if lib().pm_server_create(C.byref(self.h), int(device), int(hbm_budget_bytes), int(max_active)) != 0:
            raise PmError(lib().pm_global_last_error().decode())
        self.fems = {}

    def submit(self, fem):
        t = C.c_int64(-1)
        if lib().pm_server_submit(self.h, fem.h, C.byref(t)) != 0:
            raise PmError(lib().pm_last_error(fem.h).decode())
        self.fems[t.value] = fem
        return t.value

    def _record(self, ticket, o):
        f = self.fems[ticket]
        r = {k: getattr(o, k) for (k, _) in PmRequestStats._fields_}
        r["error"] = lib().pm_last_error(f.h).decode() if o.status != 0 else ""
        if o.status == 0:
            f.refresh_sizes()
        return r

    def poll(self, ticket):
        """The request's record once done, else None."""
        o = PmRequestStats()
        rc = lib().pm_server_poll(self.h, int(ticket), C.cast(C.pointer(o), C.c_void_p))
        if rc < 0:
            raise PmError("unknown ticket")
        return self._record(ticket, o) if rc == 0 else None

    def wait(self, ticket):
        o = PmRequestStats()
        if lib().pm_server_wait(self.h, int(ticket), C.cast(C.pointer(o), C.c_void_p)) != 0:
            raise PmError("unknown ticket")
        return self._record(ticket, o)

    def stats(self):
        v = np.zeros(2, dtype=np.int64)
        lib().pm_server_stats(self.h, v)
        return {"pass_batches": int(v[0]), "pass_jobs": int(v[1])}

    def close(self):
        if self.h:
            lib().pm_server_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
