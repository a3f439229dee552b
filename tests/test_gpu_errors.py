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


def test_general_plasticity_has_no_stiffness():
    """Settings::set_general_plasticity routes von Mises to pmsim's Elastoplastic model
    (model_stress_strain.rs:65-71), whose stiffness returns "TODO" (elastoplastic.rs:528-536): the solve
    stops at the first tangent with that text.  A linear-elastic attribute ignores the setting."""
    mesh, base, cfg, ess, nat = cases.footing_c1(n_inc=2)
    cfg.update_model_settings(1).set_general_plasticity(True)
    assert _fem_error((mesh, base, cfg, ess, nat)) == "TODO"
    f = Fem(describe(mesh, base, cfg, ess, nat), cfg)
    for call in (f.assemble_kk, f.element_pass):
        with pytest.raises(PmError, match="^TODO$"):
            call()
    with pytest.raises(PmError, match="general-plasticity"):
        f.update_secondary_values()
    f.assemble_f_int()  # the internal forces of the current state need no stress update
    mesh, base, cfg, ess, nat = _one_qua4(StressStrain.LinearElastic(1000.0, 0.25))
    f0 = Fem(describe(mesh, base, cfg, ess, nat), cfg)
    f0.solve()
    cfg.update_model_settings(1).set_general_plasticity(True).set_gp_interp_nn_max(20)
    f1 = Fem(describe(mesh, base, cfg, ess, nat), cfg)
    f1.solve()
    assert np.array_equal(f0.get_u()[0], f1.get_u()[0])


@pytest.mark.parametrize("setting, printed", [
    (lambda s: s.set_nle_beta(-1.0), "ERROR: nle_beta = -1.0 is incorrect; it must be ≥ 0.0"),
    (lambda s: s.set_gp_interp_nn_max(0), "ERROR: gp_interp_nn_max = 0 is incorrect; it must be ≥ 1"),
])
def test_model_settings_are_validated(setting, printed, capfd):
    """settings.rs:55-66 through ModelStressStrain::new (model_stress_strain.rs:53-58), for any model."""
    mesh, base, cfg, ess, nat = _one_qua4(StressStrain.LinearElastic(1000.0, 0.25))
    setting(cfg.update_model_settings(1))
    with pytest.raises(PmError) as e:
        Fem(describe(mesh, base, cfg, ess, nat), cfg)
    assert str(e.value) == "cannot allocate ModelStressStrain because settings.validate() failed"
    assert printed in capfd.readouterr().out
