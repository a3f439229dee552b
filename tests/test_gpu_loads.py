"""Distributed loads, external forces, the residual and element-level access through the C ABI (GPU).

F_ext of Qn / Qx / Qy / Qz against pmsim's known answers (src/fem/bc_distributed.rs:446-512) and the
oracle; uniform-traction blocks solved on the device against their closed-form uniform stress; the
residual R = F_int - l F_ext (solver_data.rs:111-120), y = K x, and the local K and f_int of single cells
(ElementTrait::calc_jacobian / calc_f_int) against the oracle's.
"""
import numpy as np
import pytest

import cases
import loads_cases as LC
from gpu_util import rel_err
from oracle import oracle as orc
from pmsim_amd import Nbc, describe
from pmsim_amd.solver import Fem, PmError

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nbc", [Nbc.Qn, Nbc.Qx, Nbc.Qy])
@pytest.mark.parametrize("edge", ["top", "left", "right", "bottom"])
def test_f_ext_qua8_edge_known_answers(edge, nbc):
    case = LC.qua8_edge_problem(edge, nbc)
    f = Fem(describe(*case), case[2])
    f.assemble_f_ext(0.0)
    got = f.get_f_ext().reshape(-1, 2)
    want = np.zeros_like(got)
    for p, v in LC.expected_qua8(edge, nbc).items():
        want[p] = v
    assert np.abs(got - want).max() < 1e-14


def test_f_ext_qz_on_3d_edge_known_answer():
    case = LC.hex8_edge_qz_problem()
    f = Fem(describe(*case), case[2])
    f.assemble_f_ext(0.0)
    got = f.get_f_ext().reshape(-1, 3)
    want = np.zeros_like(got)
    want[4, 2] = want[5, 2] = LC.Q / 2.0
    assert np.abs(got - want).max() < 1e-14


def test_distributed_load_errors():
    """pmsim's texts: Qn is refused on 3D edges (bc_distributed.rs:75-83), Qz has no DOF in 2D."""
    mesh, base, cfg, ess, nat = LC.hex8_edge_qz_problem()
    nat.on_facets = [(f, int(Nbc.Qn), v, fi) for (f, _, v, fi) in nat.on_facets]
    with pytest.raises(PmError, match="Qn natural boundary condition is not available for 3D edge"):
        Fem(describe(mesh, base, cfg, ess, nat), cfg)
    mesh, base, cfg, ess, nat = LC.qua8_edge_problem("top", Nbc.Qy)
    nat.on_facets = [(f, int(Nbc.Qz), v, fi) for (f, _, v, fi) in nat.on_facets]
    with pytest.raises(PmError, match="cannot find the number of a"):
        Fem(describe(mesh, base, cfg, ess, nat), cfg)


def _uniform(hist, ns, want):
    ip = hist[-1][3]
    scale = np.abs(want).max()
    return np.abs(ip[:, :ns] - np.asarray(want)[None, :]).max() / scale


@pytest.mark.parametrize("lin_sol", ["host", "pcg"])
@pytest.mark.parametrize("kind,ps", [("qua4", False), ("qua8", False), ("qua9", False), ("tri6", False), ("tri15", False),
                                     ("qua8", True)])
def test_uniform_traction_2d(kind, ps, lin_sol):
    """Qx = q on x = 2 and Qy = -q/2 on y = 1 over a roller-supported block: sigma = (q, -q/2, nu(q - q/2), 0)
    in plane strain, (q, -q/2, 0, 0) in plane stress, at every IP; F_ext equals the oracle's."""
    q = 3.0
    mesh, base, cfg, ess, nat = LC.traction_block_2d(kind, q=q, plane_stress=ps)
    cfg.set_lin_sol(lin_sol)
    d = describe(mesh, base, cfg, ess, nat)
    f = Fem(d, cfg)
    f.assemble_f_ext(0.0)
    assert np.abs(f.get_f_ext() - orc.Oracle(d).f_ext(0.0)).max() < 1e-14
    f.solve()
    szz = 0.0 if ps else LC.NU_T * (q - q / 2.0)
    assert _uniform(f.history(), f.ns, [q, -q / 2.0, szz, 0.0]) < 1e-11


@pytest.mark.parametrize("kind", ["hex8", "hex20", "tet10"])
def test_uniform_traction_3d(kind):
    """Qz = -q on the top face and Qx = q/2 on x = 1 over a roller-supported cube: sigma = (q/2, 0, -q, 0, 0, 0)."""
    q = 3.0
    mesh, base, cfg, ess, nat = LC.traction_block_3d(kind, q=q)
    cfg.set_lin_sol("pcg")
    d = describe(mesh, base, cfg, ess, nat)
    f = Fem(d, cfg)
    f.assemble_f_ext(0.0)
    assert np.abs(f.get_f_ext() - orc.Oracle(d).f_ext(0.0)).max() < 1e-14
    f.solve()
    assert _uniform(f.history(), f.ns, [q / 2.0, 0.0, -q, 0.0, 0.0, 0.0]) < 1e-11


def test_residual_apply_kk_and_element_access():
    """After one plastic update of the pressurised cylinder (SPO 7.5.1, Qn(t) loads): R = F_int - l F_ext bit for
    bit (l = 1 and 0.5), y = K x equal to the assembled CSR's product, and every cell's local K and f_int equal
    to the oracle's (1e-12)."""
    import scipy.sparse as sp
    mesh, base, cfg, ess, nat = cases.spo_751_pres_cylin()  # Qn(t) loads: a non-zero F_ext
    d = describe(mesh, base, cfg, ess, nat)
    o = orc.Oracle(d)
    f = Fem(d, cfg)
    rng = np.random.default_rng(3)
    ddu = rng.uniform(-2e-1, 2e-1, f.ndof)
    o.init_state()
    o.update_secondary(ddu)
    f.set_ddu(ddu)
    f.update_secondary_values()
    f.assemble_f_int()
    f.assemble_kk()
    f.assemble_f_ext(3.0)
    fe, fi = f.get_f_ext(), f.get_f_int()
    assert np.abs(fe - o.f_ext(3.0)).max() <= 1e-14 * np.abs(fe).max()
    for ell in (1.0, 0.5):
        f.calculate_residuals_vector(ell)
        assert np.array_equal(f.get_residual(), fi - ell * fe)
    rp, cols, vals = f.get_csr()
    x = rng.normal(size=f.ndof)
    y = f.apply_kk(x)
    assert rel_err(y, sp.csr_matrix((vals, cols, rp), shape=(f.ndof, f.ndof)) @ x) < 1e-14
    nd = d["ndim"]
    for e in range(mesh.ncell):
        ne = len(mesh.cell_points(e)) * nd
        assert rel_err(f.element_kk(e), o.element_K(e, ne)) < 1e-12
        assert rel_err(f.element_f_int(e), o.element_f(e, ne)) < 1e-12


@pytest.mark.parametrize("kind", ["qua8", "tri6", "hex20", "tet10"])
def test_global_tangent_is_exactly_symmetric(kind):
    """Element matrices are held as their upper node blocks with mirrored diagonal blocks, so every local matrix
    and K are symmetric bit for bit: pmsim's local check |K_lm - K_ml| > tol (assembly.rs:78-90) never fires."""
    from pmsim_amd import mesh as M
    from test_gpu_parity import _block_problem, _ddu
    k = M.KIND_NAMES[kind]
    mesh, d = _block_problem(k, (3, 2) if k <= M.QUA9 else (2, 2, 1))
    f = Fem(d)
    f.set_ddu(_ddu(mesh, 3e-3, 5))
    f.update_secondary_values()
    f.assemble_kk()
    import scipy.sparse as sp
    rp, cols, vals = f.get_csr()
    K = sp.csr_matrix((vals, cols, rp), shape=(f.ndof, f.ndof))
    assert (K - K.T).count_nonzero() == 0
    for e in range(mesh.ncell):
        ke = f.element_kk(e)
        assert np.array_equal(ke, ke.T)
