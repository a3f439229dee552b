"""Error paths through the C ABI on the GPU, with pmsim's texts; each condition is also checked on the
oracle in test_errors.py where the oracle models it."""
import numpy as np
import pytest

import cases
from gpu_util import rel_err
from pmsim_amd import Natural, Pbc, StressStrain, describe
from pmsim_amd.solver import Fem, PmError
from test_errors import _one_qua4

pytestmark = pytest.mark.gpu


def _fem_error(case):
    mesh, base, cfg, ess, nat = case
    with pytest.raises(PmError) as e:
        Fem(describe(mesh, base, cfg, ess, nat), cfg).solve()
    return str(e.value)


def test_lmm_required_for_non_zero_prescribed_values():
    mesh, base, cfg, ess, nat = cases.footing_c1(n_inc=2)
    cfg.set_lagrange_mult_method(False)
    assert _fem_error((mesh, base, cfg, ess, nat)) == \
        "the Lagrange multiplier method is required for non-zero prescribed values"


@pytest.mark.parametrize("lin_sol", ["host", "pcg"])
def test_newton_raphson_did_not_converge(lin_sol):
    mesh, base, cfg, ess, nat = cases.footing_c1(n_inc=2)
    cfg.set_n_max_iterations(1).set_lin_sol(lin_sol)
    assert _fem_error((mesh, base, cfg, ess, nat)) == "Newton-Raphson did not converge"


def test_too_many_step_failures():
    """solver_implicit.rs:162-165: the counter grows every step, so a limit of 2 ends the third step."""
    mesh, base, cfg, ess, nat = cases.footing_c1(n_inc=5)
    cfg.set_allowed_step_n_failure(2)
    assert _fem_error((mesh, base, cfg, ess, nat)) == "too many step failures"


def test_von_mises_parameter_errors():
    assert _fem_error(_one_qua4(StressStrain.VonMises(1000.0, 0.25, 0.0, 10.0), plane_stress=True)) == \
        "von Mises model does not work in plane-stress"
    assert _fem_error(_one_qua4(StressStrain.VonMises(1000.0, 0.25, 0.0, 1e-7))) == \
        "von Mises initial size of the yield surface must > 1e-6"


def test_nan_load_is_found():
    mesh, base, cfg, ess, _ = _one_qua4(StressStrain.LinearElastic(1000.0, 0.25))
    nat = Natural()
    nat.points([2, 3], Pbc.Fy, float("nan"))
    assert _fem_error((mesh, base, cfg, ess, nat)) == "Found NaN or Inf"


def test_point_outside_every_cell():
    """A point that no cell uses is refused by pm_initialize (the host-side FemBase refuses it too)."""
    mesh, base, cfg, ess, nat = _one_qua4(StressStrain.LinearElastic(1000.0, 0.25))
    d = describe(mesh, base, cfg, ess, nat)
    d["coords"] = np.vstack([d["coords"], [[5.0, 5.0]]])
    with pytest.raises(PmError, match="all points must belong to at least one cell"):
        Fem(d, cfg)


def test_pcg_failure_falls_back_to_host_direct():
    """A PCG that stops at its iteration limit hands the solve to the host direct solver (counted in
    the statistics); the path is the host solver's."""
    mesh, base, cfg, ess, nat = cases.footing_c1(n_inc=3)
    cfg.set_lin_sol("pcg").set_pcg(1e-12, 2)
    f = Fem(describe(mesh, base, cfg, ess, nat), cfg)
    f.solve()
    assert f.stats()["n_host_fallbacks"] > 0
    mesh, base, cfg, ess, nat = cases.footing_c1(n_inc=3)
    cfg.set_lin_sol("host")
    g = Fem(describe(mesh, base, cfg, ess, nat), cfg)
    g.solve()
    assert g.stats()["n_host_fallbacks"] == 0
    assert rel_err(f.history()[-1][1], g.history()[-1][1]) < 1e-10
