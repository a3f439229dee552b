"""Problem description: parameters, boundary conditions and configuration.

Mirrors pmsim's builder API for the solid / von Mises path so that a driver reads like the
reference's tests (``tests/test_spo_754_footing.rs:40-93``):

* ``StressStrain.VonMises/LinearElastic`` and ``ParamSolid``  -- ``src/base/parameters.rs:6-68,276-285``
* ``Essential.points/edges/faces(_fn)``                        -- ``src/base/essential.rs:38-135``
* ``Natural.points/edges/faces``                               -- ``src/base/natural.rs:50-140``
* ``Config`` (defaults and setters used on the path)           -- ``src/base/config.rs:184-231,366-664``
* ``FemBase`` (attribute -> element map, DOF count)            -- ``src/fem/fem_base.rs:25-30``

Multiplier functions ``m(t)`` are host callbacks in pmsim (``essential.rs:94-120``); here they are
evaluated once per time station into value tables that the C ABI consumes, since no host
callback may run on the device hot path.
"""
from __future__ import annotations

import dataclasses
import enum
import math
from typing import Callable, Dict, List, Optional, Sequence, Tuple

import numpy as np

from .mesh import Facet, Mesh


class Dof(enum.IntEnum):
    Ux = 0
    Uy = 1
    Uz = 2


class Pbc(enum.IntEnum):
    """Point (concentrated) loads."""
    Fx = 0
    Fy = 1
    Fz = 2


class Nbc(enum.IntEnum):
    """Distributed boundary loads (solid subset)."""
    Qn = 0
    Qx = 1
    Qy = 2
    Qz = 3


MODEL_LINEAR_ELASTIC = 0
MODEL_VON_MISES = 1


@dataclasses.dataclass(frozen=True)
class StressStrain:
    model: int
    young: float
    poisson: float
    hh: float = 0.0
    z_ini: float = 0.0

    @staticmethod
    def LinearElastic(young: float, poisson: float) -> "StressStrain":
        return StressStrain(MODEL_LINEAR_ELASTIC, young, poisson)

    @staticmethod
    def VonMises(young: float, poisson: float, hh: float, z_ini: float) -> "StressStrain":
        return StressStrain(MODEL_VON_MISES, young, poisson, hh, z_ini)


@dataclasses.dataclass(frozen=True)
class ParamSolid:
    density: float
    stress_strain: StressStrain
    ngauss: Optional[int] = None


class Elem:
    @staticmethod
    def Solid(p: ParamSolid) -> ParamSolid:
        return p


class Essential:
    """Prescribed values ``value = c * m(t)``; a later call overrides an earlier one for the same (point, dof)."""

    def __init__(self):
        self.all: Dict[Tuple[int, int], Tuple[float, Optional[int]]] = {}
        self.functions: List[Callable[[float], float]] = []

    def _put(self, pts, dof, c, f_index):
        for p in pts:
            self.all[(int(p), int(dof))] = (float(c), f_index)
        return self

    def point(self, p, dof, value):
        return self._put([p], dof, value, None)

    def points(self, pts, dof, value):
        return self._put(pts, dof, value, None)

    def edges(self, facets: Sequence[Facet], dof, value):
        return self._put([p for f in facets for p in f.points], dof, value, None)

    faces = edges

    def _fn(self, pts, dof, c, m):
        self.functions.append(m)
        return self._put(pts, dof, c, len(self.functions) - 1)

    def points_fn(self, pts, dof, c, m):
        return self._fn(pts, dof, c, m)

    def edges_fn(self, facets, dof, c, m):
        return self._fn([p for f in facets for p in f.points], dof, c, m)

    faces_fn = edges_fn

    def size(self):
        return len(self.all)

    def has_non_zero(self):
        return any(c != 0.0 for (c, _) in self.all.values())

    def tables(self, times: np.ndarray):
        """Sorted (point, dof) keys and the (n × nstation) value table at the given times."""
        keys = sorted(self.all.keys())
        tab = np.zeros((len(keys), len(times)))
        for i, k in enumerate(keys):
            c, fi = self.all[k]
            if fi is None:
                tab[i, :] = c
            else:
                f = self.functions[fi]
                tab[i, :] = [c * f(float(t)) for t in times]
        pts = np.asarray([k[0] for k in keys], dtype=np.int32)
        dofs = np.asarray([k[1] for k in keys], dtype=np.int32)
        return pts, dofs, tab


class Natural:
    """Concentrated point loads and distributed loads (Qn normal, Qx/Qy/Qz directional) on boundary
    facets (src/base/natural.rs:40-140, src/fem/bc_distributed.rs:156-236): each value is a constant
    or, with the ``_fn`` forms, ``f(t)`` evaluated at the step's time."""

    def __init__(self):
        self.at_points: List[Tuple[int, int, float, Optional[int]]] = []
        self.on_facets: List[Tuple[Facet, int, float, Optional[int]]] = []
        self.functions: List[Callable[[float], float]] = []

    def _fn_index(self, f):
        self.functions.append(f)
        return len(self.functions) - 1

    def point(self, p, pbc, value):
        self.at_points.append((int(p), int(pbc), float(value), None))
        return self

    def points(self, pts, pbc, value):
        for p in pts:
            self.point(p, pbc, value)
        return self

    def point_fn(self, p, pbc, f):
        self.at_points.append((int(p), int(pbc), 0.0, self._fn_index(f)))
        return self

    def points_fn(self, pts, pbc, f):
        i = self._fn_index(f)
        for p in pts:
            self.at_points.append((int(p), int(pbc), 0.0, i))
        return self

    def edge(self, f: Facet, nbc, value):
        self.on_facets.append((f, int(nbc), float(value), None))
        return self

    def edges(self, facets, nbc, value):
        for f in facets:
            self.edge(f, nbc, value)
        return self

    def edge_fn(self, facet: Facet, nbc, f):
        self.on_facets.append((facet, int(nbc), 0.0, self._fn_index(f)))
        return self

    def edges_fn(self, facets, nbc, f):
        i = self._fn_index(f)
        for facet in facets:
            self.on_facets.append((facet, int(nbc), 0.0, i))
        return self

    face = edge
    faces = edges
    face_fn = edge_fn
    faces_fn = edges_fn

    def values(self, value, f_index, times):
        """A load's values at the time stations (one value when it is constant)."""
        if f_index is None:
            return np.asarray([value], dtype=np.float64)
        f = self.functions[f_index]
        return np.asarray([f(float(t)) for t in times], dtype=np.float64)


class Settings:
    """Model settings of one cell attribute (``src/material/settings.rs:6-121``), obtained from
    ``Config.update_model_settings``.  ``general_plasticity`` selects, for von Mises, pmsim's general
    ``Elastoplastic`` model (``model_stress_strain.rs:65-71``), whose consistent stiffness pmsim leaves
    unimplemented (``elastoplastic.rs:528-536`` returns "TODO"): a solve then stops at its first tangent
    with that error.  The other fields only matter to that model; they are validated as pmsim does
    (``settings.rs:55-66``) when the elements are allocated."""

    def __init__(self):
        self.save_strain = False
        self.nle_enabled = False
        self.nle_beta = 0.0
        self.nle_isotropic = False
        self.general_plasticity = False
        self.gp_ode_method = "DoPri8"
        self.gp_interp_nn_max = 30
        self.gp_allow_initial_drift = False
        self.gp_save_history = False

    def set_save_strain(self, flag: bool): self.save_strain = bool(flag); return self
    def set_nle_enabled(self, flag: bool): self.nle_enabled = bool(flag); return self
    def set_nle_beta(self, value: float): self.nle_beta = float(value); return self
    def set_nle_isotropic(self, flag: bool): self.nle_isotropic = bool(flag); return self
    def set_general_plasticity(self, flag: bool): self.general_plasticity = bool(flag); return self
    def set_gp_ode_method(self, method: str): self.gp_ode_method = str(method); return self

    def set_gp_interp_nn_max(self, nn_max: int):
        if int(nn_max) < 0:
            raise ValueError("gp_interp_nn_max is an unsigned integer")
        self.gp_interp_nn_max = int(nn_max)
        return self

    def set_gp_allow_initial_drift(self, flag: bool): self.gp_allow_initial_drift = bool(flag); return self
    def set_gp_save_history(self, flag: bool): self.gp_save_history = bool(flag); return self


class Config:
    """Solver configuration (defaults of ``src/base/config.rs:184-231``)."""

    def __init__(self, mesh: Mesh):
        self.ndim = mesh.ndim
        self.linear_problem = False
        self.constant_tangent = False
        self.allowed_step_n_failure = 100
        self.lagrange_mult_method = False
        self.alt_bb_matrix_method = False
        self.symmetry_check_tolerance: Optional[float] = 1e-10
        self.axisymmetric = False
        self.plane_stress = False
        self.thickness = 1.0
        self.t_ini = 0.0
        self.t_fin = 1.0
        self.ddt: Callable[[float], float] = lambda t: 1.0
        self.ddt_out: Callable[[float], float] = lambda t: 1.0
        self.ddt_min = 1e-10
        self.n_max_time_steps = 1000
        self.n_max_iterations = 10
        self.tol_rr_abs = 1e-10
        self.tol_mdu_rel = 1e-8
        self.arc_length_method = False
        self.arc_length_psi = 1.0
        self.first_trial_loading_factor = 1.0
        # MI355X path selection (not in pmsim): linear solver and tangent precision
        self.lin_sol = "pcg"          # "pcg" (device FP64 block-Jacobi PCG) or "host" (host direct)
        self.tangent_precision = "f64"  # "f64", "f32" or "bf16" (inexact-Newton tangent)
        self.tangent_kernel = "default"  # "default", "vector", "mfma" (matrix cores), "tune" (fastest measured) or
        # "vector-wg" / "mfma-wg" (the vector / matrix-core kernel with workgroups per group of elements, never the
        # persistent Hex20 forms)
        self.use_graphs = True  # replay captured HIP graphs for element passes and PCG chunks
        self.model_settings: Dict[int, Settings] = {}
        self.pcg_small_max = -1  # DOF limit of the single-workgroup PCG (-1: library default)
        self.verbose_timesteps = False  # a line per converged step on stderr
        self.verbose_iterations = False  # a line per Newton iteration on stderr
        self.pcg_tol = 1e-12
        self.pcg_max_iter = 20000
        self.scaled_rr = 0.0           # > 0 enables ||R||inf < scaled_rr * ||F_int||inf as an extra test

    # setters (config.rs:366-664)
    def set_linear_problem(self, v=True): self.linear_problem = bool(v); return self
    def set_constant_tangent(self, v=True): self.constant_tangent = bool(v); return self
    def set_allowed_step_n_failure(self, n): self.allowed_step_n_failure = int(n); return self
    def set_lagrange_mult_method(self, v=True): self.lagrange_mult_method = bool(v); return self
    def set_alt_bb_matrix_method(self, v=True): self.alt_bb_matrix_method = bool(v); return self
    def set_symmetry_check_tolerance(self, tol): self.symmetry_check_tolerance = tol; return self
    def set_axisymmetric(self):
        if self.ndim != 2:
            raise ValueError("axisymmetric requires ndim = 2")
        self.axisymmetric = True; self.plane_stress = False; return self
    def set_plane_stress(self, thickness):
        if self.ndim != 2:
            raise ValueError("plane-stress requires ndim = 2")
        self.plane_stress = True; self.axisymmetric = False; self.thickness = float(thickness); return self
    def set_t_ini(self, t): self.t_ini = float(t); return self
    def set_t_fin(self, t): self.t_fin = float(t); return self
    def set_dt(self, f): self.ddt = f; return self
    def set_dt_out(self, f): self.ddt_out = f; return self
    def set_dt_min(self, v): self.ddt_min = float(v); return self
    def set_n_max_time_steps(self, n): self.n_max_time_steps = int(n); return self
    def set_n_max_iterations(self, n): self.n_max_iterations = int(n); return self
    def set_tol_rr_abs(self, v): self.tol_rr_abs = float(v); return self
    def set_tol_mdu_rel(self, v): self.tol_mdu_rel = float(v); return self
    def set_arc_length_method(self, v=True): self.arc_length_method = bool(v); return self
    def set_arc_length_psi(self, psi): self.arc_length_psi = float(psi); return self
    def set_first_trial_loading_factor(self, ell0): self.first_trial_loading_factor = float(ell0); return self

    def validate(self) -> Optional[str]:
        """The arc-length checks of ``config.rs:255-265`` (None when valid)."""
        if self.arc_length_psi < 0.0 or self.arc_length_psi > 1.0:
            return f"arc_length_psi = {self.arc_length_psi!r} is incorrect; it must be 0.0 ≤ ψ ≤ 1.0"
        if abs(self.first_trial_loading_factor) < 1e-12:
            return (f"absolute first trial loading factor |ℓ₀| = {self.first_trial_loading_factor!r} is incorrect; "
                    "it must be ≥ 1e-12")
        return None

    def set_incremental(self, n_station: int):
        """``config.rs:538-546``: t_ini = 0, Δt = 1, t_fin = n_station - 1."""
        self.set_t_ini(0.0).set_dt(lambda t: 1.0).set_dt_out(lambda t: 1.0)
        return self.set_t_fin(float(n_station - 1) if n_station > 1 else 0.0)

    # MI355X-specific knobs
    def set_lin_sol(self, kind: str): self.lin_sol = kind; return self
    def set_tangent_precision(self, p: str): self.tangent_precision = p; return self

    def update_model_settings(self, cell_attribute: int) -> Settings:
        """The attribute's model settings, created with the defaults on first use (config.rs:515-517)."""
        return self.model_settings.setdefault(int(cell_attribute), Settings())
    def set_tangent_kernel(self, k: str): self.tangent_kernel = k; return self
    def set_use_graphs(self, v: bool): self.use_graphs = bool(v); return self
    def set_pcg_small_max(self, n: int): self.pcg_small_max = int(n); return self
    def set_verbose(self, timesteps: bool, iterations: bool): self.verbose_timesteps = timesteps; self.verbose_iterations = iterations; return self
    def set_pcg(self, tol=1e-12, max_iter=20000): self.pcg_tol = float(tol); self.pcg_max_iter = int(max_iter); return self
    def set_scaled_rr(self, v): self.scaled_rr = float(v); return self

    def stations(self) -> Tuple[float, np.ndarray]:
        """Constant Δt (evaluated at t_ini) and the time stations t_k = t_ini + k Δt covering [t_ini, t_fin]."""
        dt = float(self.ddt(self.t_ini))
        if dt < self.ddt_min:
            raise ValueError("Δt is smaller than the allowed minimum")
        n = int(math.floor((self.t_fin - self.t_ini) / dt + 1e-9)) + 1
        return dt, self.t_ini + dt * np.arange(max(n, 1))


class FemBase:
    """Attribute -> ParamSolid map and DOF numbering (pure solid: eq = ndim*p + d)."""

    def __init__(self, mesh: Mesh, amap: Sequence[Tuple[int, ParamSolid]]):
        self.mesh = mesh
        self.amap: Dict[int, ParamSolid] = {int(a): p for (a, p) in amap}
        for a in np.unique(mesh.attrs):
            if int(a) not in self.amap:
                raise ValueError("cannot find attribute in the element map")
        used = np.zeros(mesh.npoint, dtype=bool)
        used[mesh.cell_pts] = True
        if not used.all():
            raise ValueError("all points must belong to at least one cell")
        self.ndof = mesh.ndim * mesh.npoint

    def eq(self, point: int, dof: int) -> int:
        if dof >= self.mesh.ndim:
            raise ValueError("cannot find (PointId, Dof) in the DOF map")
        return self.mesh.ndim * int(point) + int(dof)


def describe(mesh: Mesh, base: FemBase, config: Config, essential: Essential, natural: Natural) -> dict:
    """Flat description of a problem (plain arrays) consumed by the C ABI and by the oracle."""
    dt, times = config.stations()
    pts, dofs, tab = essential.tables(times)
    params = {}
    for a, p in base.amap.items():
        ss = p.stress_strain
        params[a] = (ss.model, ss.young, ss.poisson, ss.hh, ss.z_ini, int(p.ngauss or 0))
    pl = [(p, d, natural.values(v, fi, times)) for (p, d, v, fi) in natural.at_points]
    nst_pl = max([len(x[2]) for x in pl], default=1)
    pl_table = np.zeros((len(pl), nst_pl))
    for i, (_, _, vals) in enumerate(pl):
        pl_table[i, :] = vals  # a constant fills its row
    facets = []
    for (f, nbc, v, fi) in natural.on_facets:
        if nbc not in (Nbc.Qn, Nbc.Qx, Nbc.Qy, Nbc.Qz):
            raise ValueError("unknown distributed load")
        facets.append((f.kind, tuple(f.points), f.cell, natural.values(v, fi, times), int(nbc)))
    return dict(
        ndim=mesh.ndim, coords=mesh.coords, kinds=mesh.kinds, attrs=mesh.attrs,
        cell_ptr=mesh.cell_ptr, cell_pts=mesh.cell_pts, params=params,
        model_settings={a: (int(m.general_plasticity), int(m.gp_interp_nn_max), float(m.nle_beta))
                        for a, m in config.model_settings.items()},
        densities={a: float(p.density) for a, p in base.amap.items()},
        axisym=config.axisymmetric, plane_stress=config.plane_stress, thickness=config.thickness,
        lmm=config.lagrange_mult_method, linear_problem=config.linear_problem,
        t_ini=config.t_ini, t_fin=config.t_fin, dt=dt, n_max_iterations=config.n_max_iterations,
        tol_rr_abs=config.tol_rr_abs, tol_mdu_rel=config.tol_mdu_rel,
        arc_length_method=config.arc_length_method, arc_length_psi=config.arc_length_psi,
        first_trial_loading_factor=config.first_trial_loading_factor,
        ess_points=pts, ess_dofs=dofs, ess_table=tab,
        pl_points=np.asarray([x[0] for x in pl], dtype=np.int32),
        pl_dofs=np.asarray([x[1] for x in pl], dtype=np.int32),
        pl_table=pl_table,
        facets=facets,
    )
