"""Benchmark workloads (mesh + material + boundary conditions + increments)."""
import numpy as np

from .generators import hex20_block, qua8_block, mixed_2d
from .mesh import At, Features
from .problem import Config, Dof, Elem, Essential, FemBase, Natural, ParamSolid, StressStrain


def hex20_cube(nx=64, ny=64, nz=32, n_inc=10, uz=-0.01, seed=20261203, jitter=0.05, stations=None):
    """Unit cube of nx x ny x nz Hex20 (27 IP), corner jitter ±5% of h (seeded), von Mises E=2e5 nu=0.3
    z0=250 H=1e3; bottom face fixed, top face uz ramping linearly to `uz` in n_inc increments, or taking
    the values `stations` (station 0 first) when given."""
    mesh = hex20_block(nx, ny, nz, 1.0, 1.0, 1.0, seed=seed, jitter=jitter)
    feat = Features(mesh)
    bottom = feat.search_faces(At.Z(0.0))
    top = feat.search_faces(At.Z(1.0))
    base = FemBase(mesh, [(1, Elem.Solid(ParamSolid(1.0, StressStrain.VonMises(2e5, 0.3, 1e3, 250.0), 27)))])
    ess = Essential()
    ess.faces(bottom, Dof.Ux, 0.0).faces(bottom, Dof.Uy, 0.0).faces(bottom, Dof.Uz, 0.0)
    cfg = Config(mesh)
    if stations is None:
        ess.faces_fn(top, Dof.Uz, uz, lambda t: t / n_inc)
        cfg.set_incremental(n_inc + 1)
    else:
        uzs = [float(v) for v in stations]
        ess.faces_fn(top, Dof.Uz, 1.0, lambda t: uzs[int(t)])
        cfg.set_incremental(len(uzs))
    cfg.set_lagrange_mult_method(True).set_n_max_iterations(20)
    return mesh, base, cfg, ess, Natural()


def strip_footing(nx=16, ny=8, ngauss=9, n_inc=20, young=1e4, poisson=0.3, z_ini=200.0, width=1.0, uy=-0.05,
                  lx=10.0, ly=5.0):
    """Plane-strain strip footing: lx x ly domain of nx x ny Qua8, von Mises; bottom fixed, lateral ux = 0,
    uy ramping to `uy` over the top edges within x <= width, in n_inc increments."""
    mesh = qua8_block(nx, ny, lx, ly)
    feat = Features(mesh)
    left = feat.search_edges(At.X(0.0)); right = feat.search_edges(At.X(lx)); bottom = feat.search_edges(At.Y(0.0))
    foot = feat.search_edges(At.Y(ly), lambda x: x[0] <= width + 1e-12)
    base = FemBase(mesh, [(1, Elem.Solid(ParamSolid(1.0, StressStrain.VonMises(young, poisson, 0.0, z_ini), ngauss)))])
    ess = Essential()
    ess.edges(left, Dof.Ux, 0.0).edges(right, Dof.Ux, 0.0).edges(bottom, Dof.Ux, 0.0).edges(bottom, Dof.Uy, 0.0)
    ess.edges_fn(foot, Dof.Uy, uy, lambda t: t / n_inc)
    cfg = Config(mesh)
    cfg.set_lagrange_mult_method(True).set_incremental(n_inc + 1).set_n_max_iterations(20)
    return mesh, base, cfg, ess, Natural()


def mixed_compression(n_elements=200_000, n_inc=8, uy=-0.02, lx=100.0, ly=100.0, seed=20261203, jitter=0.1,
                      young=1e4, poisson=0.25, z_ini=100.0):
    """Mixed 2D mesh over lx x ly (30% Qua8, 25% Qua4, 20% Tri6, 15% Qua9, 10% Tri15 by element count,
    seeded, node jitter), plane strain von Mises E=1e4 nu=0.25 z0=100 H=0; bottom edges fixed, top edges
    uy ramping to `uy` in n_inc increments.  Each kind has its own attribute so that its integration rule
    is given explicitly (pmsim's ngauss is per attribute), with point sets that pmsim's own data fix:
    Qua8 and Qua9 3x3 and Qua4 2x2 Gauss-Legendre, Tri6 the 3-point interior rule of pmsim's Tri3 state
    files, Tri15 the 12-point rule of Smith & Griffiths example 5.7 (tests/test_solid_smith_5d7_tri15_plane_strain.rs:67)."""
    from .mesh import QUA4, QUA8, QUA9, TRI6, TRI15
    mesh = mixed_2d(n_elements, lx, ly, seed=seed, jitter=jitter)
    attr_ng = {QUA8: (1, 9), QUA4: (2, 4), TRI6: (3, 3), QUA9: (4, 9), TRI15: (5, 12)}
    mesh.attrs = np.asarray([attr_ng[int(k)][0] for k in mesh.kinds], dtype=np.int32)
    feat = Features(mesh)
    bottom = feat.search_edges(At.Y(0.0))
    top = feat.search_edges(At.Y(ly))
    ss = StressStrain.VonMises(young, poisson, 0.0, z_ini)
    base = FemBase(mesh, [(a, Elem.Solid(ParamSolid(1.0, ss, ng))) for a, ng in sorted(attr_ng.values())])
    ess = Essential()
    ess.edges(bottom, Dof.Ux, 0.0).edges(bottom, Dof.Uy, 0.0)
    ess.edges_fn(top, Dof.Uy, uy, lambda t: t / n_inc)
    cfg = Config(mesh)
    cfg.set_lagrange_mult_method(True).set_incremental(n_inc + 1).set_n_max_iterations(20)
    return mesh, base, cfg, ess, Natural()


def footing_request(rng, nx_range=(8, 256), ny_range=(4, 128)):
    """One parametric strip-footing solve (10 m x 5 m, Qua8, 3x3 IP): nx, ny log-uniform in their ranges,
    E in [5e3, 5e4], nu in [0.2, 0.4], z0 in [50, 500], footing width B in [0.5, 3] m, 5-40 increments
    of a footing settlement of 0.05 m.  Returns (problem, parameters)."""
    def logu(lo, hi):
        return int(round(np.exp(rng.uniform(np.log(lo), np.log(hi)))))
    p = dict(nx=logu(*nx_range), ny=logu(*ny_range), young=float(rng.uniform(5e3, 5e4)),
             poisson=float(rng.uniform(0.2, 0.4)), z_ini=float(rng.uniform(50.0, 500.0)),
             width=float(rng.uniform(0.5, 3.0)), n_inc=int(rng.integers(5, 41)))
    # the footing covers at least one element edge (10/nx), so coarse meshes still carry the load
    p["width_used"] = max(p["width"], 10.0 / p["nx"])
    prob = strip_footing(p["nx"], p["ny"], ngauss=9, n_inc=p["n_inc"], young=p["young"], poisson=p["poisson"],
                         z_ini=p["z_ini"], width=p["width_used"], uy=-0.05)
    return prob, p


def poisson_arrivals(rng, n, rate):
    """Arrival times (s) of n requests from a Poisson process of the given rate (requests/s)."""
    return np.cumsum(rng.exponential(1.0 / rate, n))
