"""CPU-side checks of the product boundary: the C-ABI library loads and exports every entry point
declared in include/pmsim_amd.h (no compute call: there is no GPU here), the flat problem
description matches pmsim's builder semantics, and the mesh generators are sound."""
import os
import re

import numpy as np
import pytest

import cases
from pmsim_amd import At, Config, Essential, Features, describe
from pmsim_amd import generators as G
from pmsim_amd import mesh as M

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "pmsim_amd.h")).read()
    return sorted(set(re.findall(r"\b(pm_[a-z_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    from pmsim_amd import solver
    L = solver.lib()
    names = _declared()
    assert len(names) >= 30
    for n in names:
        assert hasattr(L, n), n
    assert sorted(solver.EXPORTED) == names


def test_library_errors_without_device():
    """pm_fem_create validates input without touching a device; pm_initialize fails cleanly without one."""
    import ctypes as C
    from pmsim_amd import solver
    L = solver.lib()
    h = C.c_void_p()
    X = np.zeros(2)
    z = np.zeros(1, dtype=np.int32)
    rc = L.pm_fem_create(C.byref(h), 2, 1, X, 0, z, z, np.zeros(1, dtype=np.int64), z)
    assert rc != 0 and L.pm_global_last_error().decode() == "there are no cells in the mesh"


def test_c_abi_rejects_null_and_bad_arguments():
    """Every entry point checks its handle and required pointers instead of dereferencing them; bad
    tables and cell pointers are refused with a message (all without a device)."""
    import ctypes as C
    from pmsim_amd import solver
    solver.lib()
    R = C.CDLL(solver.LIB_PATH)  # a second view of the same library, without argument declarations
    R.pm_global_last_error.restype = C.c_char_p
    R.pm_last_error.restype = C.c_char_p
    R.pm_last_error.argtypes = [C.c_void_p]
    gerr = lambda: R.pm_global_last_error().decode()
    d = C.c_double
    assert R.pm_set_param_solid(None, 1, 0, d(1.0), d(0.25), d(0.0), d(0.0), 0) == 1 and gerr() == "null handle"
    for call in (lambda: R.pm_set_idealization(None, 0, 0, d(1.0)), lambda: R.pm_set_config(None, None),
                 lambda: R.pm_set_essential(None, C.c_int64(0), None, None, None, 1),
                 lambda: R.pm_set_point_loads(None, C.c_int64(0), None, None, None, 1),
                 lambda: R.pm_add_facet_qn(None, 0, 0, None, C.c_int64(0), None, 1),
                 lambda: R.pm_solve(None, None), lambda: R.pm_get_u(None, None, None),
                 lambda: R.pm_get_ip_state(None, None), lambda: R.pm_get_stats(None, None),
                 lambda: R.pm_history_count(None, None), lambda: R.pm_get_history(None, 0, None, None, None, None),
                 lambda: R.pm_get_csr(None, None, None, None), lambda: R.pm_sizes(None, None),
                 lambda: R.pm_time_element_pass(None, 1, None), lambda: R.pm_linear_solve(None, None, None, None)):
        assert call() == 1 and gerr() == "null handle"
    assert R.pm_fem_create(None, 2, C.c_int64(1), None, C.c_int64(1), None, None, None, None) == 1
    assert gerr() == "null output handle"
    h = C.c_void_p()
    X = np.zeros(8)
    kinds = np.full(1, M.QUA4, dtype=np.int32)
    attrs = np.ones(1, dtype=np.int32)
    pts = np.arange(4, dtype=np.int32)
    bad_ptr = np.array([1, 4], dtype=np.int64)
    assert R.pm_fem_create(C.byref(h), 2, C.c_int64(4), X.ctypes, C.c_int64(1), kinds.ctypes, attrs.ctypes,
                           bad_ptr.ctypes, pts.ctypes) == 1
    assert gerr() == "cell pointers must start at 0 and not decrease"
    ptr = np.array([0, 4], dtype=np.int64)
    assert R.pm_fem_create(C.byref(h), 2, C.c_int64(4), X.ctypes, C.c_int64(1), kinds.ctypes, attrs.ctypes,
                           ptr.ctypes, pts.ctypes) == 0
    try:
        p1, d1, t1 = np.zeros(1, dtype=np.int32), np.zeros(1, dtype=np.int32), np.zeros(1)
        assert R.pm_set_essential(h, C.c_int64(1), p1.ctypes, d1.ctypes, t1.ctypes, 0) == 1
        assert R.pm_last_error(h).decode() == "an essential table needs at least one station"
        assert R.pm_set_config(h, None) == 1 and R.pm_last_error(h).decode() == "null configuration"
        assert R.pm_solve(h, None) == 1 and R.pm_last_error(h).decode() == "pm_initialize has not been called"
    finally:
        R.pm_fem_destroy(h)


def test_essential_override_and_tables():
    mesh, base, cfg, ess, nat = cases.spo_754_footing()
    d = describe(mesh, base, cfg, ess, nat)
    # bottom-left corner: Ux from the left edge and Uy from the bottom edge; footing overrides nothing at x=0,y=0
    pts = set(zip(d["ess_points"].tolist(), d["ess_dofs"].tolist()))
    assert len(pts) == len(d["ess_points"])
    assert d["ess_table"].shape[1] == 15
    k = [i for i, (p, dof) in enumerate(zip(d["ess_points"], d["ess_dofs"])) if dof == 1 and mesh.coords[p][1] == 500.0]
    assert np.allclose(d["ess_table"][k[0]], [0.0, -0.01, -0.015, -0.02, -0.025, -0.035, -0.045, -0.055, -0.065, -0.075,
                                              -0.08, -0.09, -0.11, -0.14, -0.2])


def test_natural_load_tables():
    """Natural loads (src/base/natural.rs): a constant keeps one value, an ``_fn`` load takes f(t) at every
    station; point and circle selectors find the spo_751 ring's inner arc and the plate's corner."""
    mesh, base, cfg, ess, nat = cases.spo_751_pres_cylin()
    d = describe(mesh, base, cfg, ess, nat)
    assert len(d["facets"]) == 3  # three Qua4 edges on the inner circle r = 100
    for (_, pts, _, vals, nbc) in d["facets"]:
        assert nbc == 0  # Qn
        assert np.allclose(np.hypot(*mesh.coords[list(pts)].T), 100.0)
        assert np.array_equal(vals, -np.asarray(cases.P_751))
    nat.point(0, 0, 2.5).points_fn([1, 2], 1, lambda t: 10.0 * t)
    d = describe(mesh, base, cfg, ess, nat)
    assert d["pl_table"].shape == (3, 6)
    assert np.array_equal(d["pl_table"][0], np.full(6, 2.5))
    assert np.array_equal(d["pl_table"][2], 10.0 * np.arange(6.0))
    plate = cases.spo_753_circ_plate()[0]
    assert Features(plate).search_point_ids(At.XY(10.0, 0.0)) == [7]  # "7 0 10.0 0.0" in the .msh


def test_footing_edge_selection_rule():
    """Edges are selected when all their points pass the filter: nodes x in {0, 0.3125, 0.625} for the
    16x8 footing (the 0.9375 mid-node belongs to an edge that leaves x <= 1)."""
    mesh, base, cfg, ess, nat = cases.footing_c1()
    d = describe(mesh, base, cfg, ess, nat)
    top = [p for p, dof in zip(d["ess_points"], d["ess_dofs"]) if dof == 1 and abs(mesh.coords[p][1] - 5.0) < 1e-12]
    xs = sorted(mesh.coords[top][:, 0].tolist())
    assert xs == [0.0, 0.3125, 0.625]


def test_hex20_block_counts():
    m = G.hex20_block(4, 3, 2, seed=1, jitter=0.05)
    assert m.ncell == 24
    nodes = (5 * 4 * 3) + 4 * 4 * 3 + 5 * 3 * 3 + 5 * 4 * 2
    assert m.npoint == nodes
    # jittered corners keep a positive Jacobian at the centre of every element
    from oracle import oracle as orc
    for e in range(m.ncell):
        pts = m.cell_points(e)
        N, dN = orc.shape(M.HEX20, [0.0, 0.0, 0.0])
        J = m.coords[pts].T @ dN[:60].reshape(20, 3)
        assert np.linalg.det(J) > 0


def test_c2_sizes():
    """64x64x32 Hex20: 131,072 cells and 549,185 points (1,647,555 DOFs)."""
    nx, ny, nz = 64, 64, 32
    npts = (nx + 1) * (ny + 1) * (nz + 1) + nx * (ny + 1) * (nz + 1) + (nx + 1) * ny * (nz + 1) + (nx + 1) * (ny + 1) * nz
    assert npts == 549185 and 3 * npts == 1647555


@pytest.mark.parametrize("kind", [M.TRI3, M.TRI15, M.QUA9, M.TET10])
def test_block_generator_positive(kind):
    nd = 2 if kind <= M.QUA9 else 3
    m = G.block(kind, (3, 2) if nd == 2 else (2, 2, 2), (1.0,) * nd, seed=3, jitter=0.1)
    from oracle import oracle as orc
    for e in range(m.ncell):
        pts = m.cell_points(e)
        xw = orc.gauss(kind)
        for q in range(len(xw)):
            N, dN = orc.shape(kind, xw[q, :nd])
            J = m.coords[pts].T @ dN[: len(pts) * nd].reshape(len(pts), nd)
            assert np.linalg.det(J) > 0


def test_mixed_2d_fractions():
    m = G.mixed_2d(n_elements=4000, seed=5)
    frac = {k: np.mean(m.kinds == k) for k in (M.QUA8, M.QUA4, M.TRI6, M.QUA9, M.TRI15)}
    assert abs(m.ncell - 4000) < 200
    for k, want in [(M.QUA8, .30), (M.QUA4, .25), (M.TRI6, .20), (M.QUA9, .15), (M.TRI15, .10)]:
        assert abs(frac[k] - want) < 0.04


def test_model_settings_in_the_description():
    """Config::update_model_settings (config.rs:515-517): created with settings.rs's defaults on first use,
    the same object afterwards; describe() hands each attribute's settings to the C ABI."""
    from pmsim_amd import Settings
    mesh, base, cfg, ess, nat = cases.footing_c1(n_inc=2)
    d0 = describe(mesh, base, cfg, ess, nat)
    assert d0["model_settings"] == {}
    s = cfg.update_model_settings(1)
    assert isinstance(s, Settings) and s.gp_interp_nn_max == 30 and s.gp_ode_method == "DoPri8" and not s.general_plasticity
    assert cfg.update_model_settings(1).set_general_plasticity(True).set_gp_interp_nn_max(20) is s
    assert describe(mesh, base, cfg, ess, nat)["model_settings"] == {1: (1, 20, 0.0)}
    with pytest.raises(ValueError):
        s.set_gp_interp_nn_max(-1)
