"""GPU parity of the matrix-core (MFMA) tangent kernels against the oracle.

FP64 on the matrix cores agrees with the oracle's dense B^T D B to 1e-12 relative (to the largest |K|),
like the vector kernel.  The reduced-precision tangents are inexact-Newton Jacobians: FP32 operands and
accumulation agree to 2e-5 relative, BF16 operands with FP32 accumulation to 3e-2 relative (8-bit
significands); the residual, the stress update and the linear solve stay FP64, so Newton converges to
the FP64 solution within its own tolerance.
"""
import numpy as np
import pytest

import cases
from gpu_util import rel_err
from oracle import oracle as orc
from pmsim_amd import Config, Elem, Essential, FemBase, Natural, ParamSolid, StressStrain, describe
from pmsim_amd import generators as G
from pmsim_amd import mesh as M
from pmsim_amd.solver import Fem

pytestmark = pytest.mark.gpu
KINDS = [M.TRI3, M.TRI6, M.TRI10, M.TRI15, M.QUA4, M.QUA8, M.QUA9, M.TET4, M.TET10, M.HEX8, M.HEX20]
TOL = {"f64": 1e-12, "f32": 2e-5, "bf16": 3e-2}


def _problem(kind, n, axisym=False):
    from test_gpu_parity import PINNED_NGAUSS
    nd = 2 if kind <= M.QUA9 else 3
    mesh = G.block(kind, n, (1.0,) * nd, seed=11, jitter=0.1)
    if axisym:
        mesh.coords[:, 0] += 0.5
    base = FemBase(mesh, [(1, Elem.Solid(ParamSolid(1.0, StressStrain.VonMises(1000.0, 0.3, 60.0, 1.0), PINNED_NGAUSS.get(kind))))])
    cfg = Config(mesh)
    if axisym:
        cfg.set_axisymmetric()
    return mesh, cfg, describe(mesh, base, cfg, Essential(), Natural())


def _ddu(mesh, scale, seed):
    rng = np.random.default_rng(seed)
    nd = mesh.ndim
    A = rng.uniform(-scale, scale, (nd, nd))
    return ((mesh.coords @ A.T) + rng.uniform(-scale / 5, scale / 5, mesh.coords.shape)).ravel()


def _plastic_pair(kind, n, cfg_fn, axisym=False):
    """Oracle and HIP handle after the same partly plastic update (consistent-tangent branch active)."""
    mesh, cfg, d = _problem(kind, n, axisym)
    cfg_fn(cfg)
    o = orc.Oracle(d)
    f = Fem(d, cfg)
    o.init_state()
    ddu = _ddu(mesh, 3e-3, 101)
    o.update_secondary(ddu)
    f.set_ddu(ddu)
    f.update_secondary_values()
    assert o.get_ip_state()[:, o.ns + 2].min() == 0.0  # some IPs plastic
    return o, f


@pytest.mark.parametrize("prec", ["f64", "f32", "bf16"])
@pytest.mark.parametrize("kind", KINDS)
def test_mfma_tangent_parity(kind, prec):
    n = (3, 2) if kind <= M.QUA9 else (2, 2, 1)
    o, f = _plastic_pair(kind, n, lambda c: c.set_tangent_kernel("mfma").set_tangent_precision(prec))
    f.assemble_kk()
    rp_o, c_o, v_o = o.kk_csr()
    rp_f, c_f, v_f = f.get_csr()
    assert np.array_equal(rp_o, rp_f) and np.array_equal(c_o, c_f), "CSR pattern differs"
    assert rel_err(v_f, v_o) < TOL[prec]


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("kind", [M.QUA4, M.QUA8, M.TRI6])
def test_mfma_tangent_axisymmetric(kind, prec):
    o, f = _plastic_pair(kind, (2, 2), lambda c: c.set_tangent_kernel("mfma").set_tangent_precision(prec), axisym=True)
    f.assemble_kk()
    assert rel_err(f.get_csr()[2], o.kk_csr()[2]) < TOL[prec]


def test_tangent_kernels_agree_and_tuner_picks_one():
    """Vector, matrix-core and tuned FP64 kernels give the same K to rounding; each is deterministic."""
    vals = {}
    for k in ("vector", "mfma", "tune"):
        o, f = _plastic_pair(M.HEX20, (3, 2, 2), lambda c, k=k: c.set_tangent_kernel(k))
        f.assemble_kk()
        v = f.get_csr()[2]
        f.assemble_kk()
        assert np.array_equal(f.get_csr()[2], v)
        vals[k] = v
    assert rel_err(vals["mfma"], vals["vector"]) < 1e-13
    assert rel_err(vals["tune"], vals["vector"]) < 1e-13


@pytest.mark.parametrize("prec", ["f32", "bf16"])
def test_inexact_newton_footing(prec):
    """Strip footing 16x8 Qua8 with a reduced-precision tangent: converges in every increment to the
    FP64 solution (the oracle's) within the 1e-8 relative displacement bar; FP64 residuals keep the fixed
    point exact.  Measured on MI355X: FP32 1.9e-12 (66 Newton iterations against the oracle's 52), BF16
    3.4e-9 (102)."""
    mesh, base, cfg, ess, nat = cases.footing_c1()
    cfg.set_tangent_precision(prec).set_n_max_iterations(60)
    d = describe(mesh, base, cfg, ess, nat)
    f = Fem(d, cfg)
    f.solve()
    hist = f.history()
    uo, ipo, _ = orc.Oracle(d).solve()
    assert len(hist) == len(uo)
    for k in range(len(hist)):
        assert rel_err(hist[k][1], uo[k][: f.ndof]) < 1e-8
    assert f.stats()["n_iterations"] > 0
