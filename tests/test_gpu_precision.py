"""Reduced-precision tangents (inexact Newton) on Hex20 against the FP64 oracle.

The precision forms of the 96x96x48 Hex20 cube, at sizes the oracle solves in seconds: the same jittered
unit cube and material (von Mises E=2e5, nu=0.3, z0=250, H=1e3, 27 IPs), bottom fixed and the top
pressed down in 10 increments of uz = -0.001 (the loading of the 64x64x32 cube, uz -> -0.01 in 10
increments), so most IPs end plastic.  The element tangents are held in FP32, as the vector kernel's FP64
products rounded when stored (the FP32 default) or formed on the matrix cores from FP32 operands (`mfma`),
or formed on the matrix cores from BF16 operands; every matrix-core form accumulates in FP32.  The residual, the return map, the PCG vectors and the convergence tests stay FP64,
so Newton converges to the FP64 fixed point.

Measured on MI355X (largest relative difference to the oracle over all increments, Newton iterations
against the oracle's 42 / 43):

| mesh  | tangent             | u       | sigma   | epbar   | Newton |
|-------|---------------------|---------|---------|---------|--------|
| 4x4x2 | FP32 vector         | 1.8e-11 | 2.6e-11 | 2.0e-11 | 43     |
| 4x4x2 | FP32 matrix cores   | 3.7e-11 | 6.6e-11 | 3.3e-11 | 43     |
| 4x4x2 | BF16                | 2.0e-9  | 2.2e-9  | 4.4e-9  | 61     |
| 6x6x3 | FP32 vector         | 5.3e-11 | 6.2e-11 | 5.4e-11 | 44     |
| 6x6x3 | FP32 matrix cores   | 8.1e-11 | 1.2e-10 | 1.4e-10 | 44     |
| 6x6x3 | BF16                | 1.2e-9  | 3.5e-9  | 2.4e-9  | 62     |

All meet the 1e-8 relative displacement bar; the tests hold u, sigma and epbar to 1e-8, the plastic
flags to equality, FP32 to at most two extra Newton iterations and BF16 to at most twice the oracle's.
"""
import numpy as np
import pytest

import cases
from gpu_util import rel_err
from oracle import oracle as orc
from pmsim_amd import describe
from pmsim_amd.solver import Fem

pytestmark = pytest.mark.gpu
BAR = 1e-8


@pytest.mark.parametrize("n", [(4, 4, 2), (6, 6, 3)])
@pytest.mark.parametrize("prec", ["f32", "f32-mfma", "bf16"])
def test_hex20_inexact_newton_vs_oracle(prec, n):
    mesh, base, cfg, ess, nat = cases.hex20_cube(*n, n_inc=10, uz=-0.01)
    cfg.set_n_max_iterations(60)
    d = describe(mesh, base, cfg, ess, nat)
    uo, ipo, its = orc.Oracle(d).solve()
    if prec == "f32-mfma":  # the FP32 tangent on the matrix cores (FP32 operands and accumulation)
        cfg.set_tangent_kernel("mfma")
    cfg.set_tangent_precision(prec.split("-")[0]).set_lin_sol("pcg")
    f = Fem(describe(mesh, base, cfg, ess, nat), cfg)
    f.solve()
    hist = f.history()
    ns = f.ns
    assert len(hist) == len(uo) == 11
    eu = es = ee = 0.0
    for k, (t, u, lam, ip) in enumerate(hist):
        eu = max(eu, rel_err(u, uo[k][: f.ndof]))
        es = max(es, rel_err(ip[:, :ns], ipo[k][:, :ns]))
        if k > 0:
            ee = max(ee, rel_err(ip[:, ns + 3], ipo[k][:, ns + 3]))
        assert np.array_equal(ip[:, ns + 2], ipo[k][:, ns + 2]), f"plastic flags differ at output {k}"
    newton, newton_o = f.stats()["n_iterations"], int(np.abs(its).sum())
    print(f"{prec} {n}: u {eu:.2e} sigma {es:.2e} epbar {ee:.2e} Newton {newton} (oracle {newton_o})", flush=True)
    assert eu < BAR and es < BAR and ee < BAR
    assert np.mean(hist[-1][3][:, ns + 2] == 0.0) > 0.9  # plastic almost everywhere at the end
    if prec.startswith("f32"):
        assert newton <= newton_o + 2
    else:
        assert newton_o < newton <= 2 * newton_o  # linear convergence of the rounded Jacobian costs iterations
