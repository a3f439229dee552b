"""The K gather on meshes where one node is shared by more than 64 cells.

The gather gives every wavefront a range of whole CSR blocks with at most 64 contributions (one lane per
contribution); a block with more contributions has a range of its own and is summed in rounds of 64.  A
fan of n triangles around one centre point makes the centre's diagonal block such a block (n
contributions: two, three and four rounds for n = 70, 130, 200), while the blocks between the centre and
the rim (two contributions each) and the rim's own blocks pack several to a range.  K is checked against
the oracle's dense B^T D B assembled per element (1e-12 relative to the largest entry, FP64; the FP32
tangent at its 2e-5 bar) after a partly plastic update, and two assemblies are bitwise equal.
"""
import math

import numpy as np
import pytest

from gpu_util import pair, rel_err
from pmsim_amd import Config, Elem, Essential, FemBase, Natural, ParamSolid, StressStrain, describe
from pmsim_amd import mesh as M

pytestmark = pytest.mark.gpu


def _fan(n, seed=3):
    """n Tri3 cells (centre, rim i, rim i+1) around the origin, rim radii jittered (counter-clockwise)."""
    rng = np.random.default_rng(seed)
    ang = 2.0 * math.pi * np.arange(n) / n
    rad = 1.0 + rng.uniform(-0.05, 0.05, n)
    coords = np.vstack([[0.0, 0.0], np.column_stack([rad * np.cos(ang), rad * np.sin(ang)])])
    cells = [(1, M.TRI3, (0, 1 + i, 1 + (i + 1) % n)) for i in range(n)]
    return M.Mesh(2, coords, cells)


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("n", [70, 130, 200])
def test_gather_node_shared_by_many_cells(n, prec):
    mesh = _fan(n)
    base = FemBase(mesh, [(1, Elem.Solid(ParamSolid(1.0, StressStrain.VonMises(1000.0, 0.3, 60.0, 30.0), 1)))])
    cfg = Config(mesh).set_tangent_precision(prec)
    d = describe(mesh, base, cfg, Essential(), Natural())
    o, f = pair(d, cfg)
    o.init_state()
    rng = np.random.default_rng(n)
    X = mesh.coords  # strains growing with x: about a third of the cells yield
    ddu = (np.column_stack([0.03 * X[:, 0] * X[:, 0], -0.015 * X[:, 0] * X[:, 1]]) + rng.uniform(-1e-4, 1e-4, X.shape)).ravel()
    o.update_secondary(ddu)
    f.set_ddu(ddu)
    f.update_secondary_values()
    plastic = o.get_ip_state()[:, o.ns + 2] == 0.0
    assert plastic.any() and not plastic.all()
    f.assemble_kk()
    rp_o, c_o, v_o = o.kk_csr()
    rp_f, c_f, v_f = f.get_csr()
    assert np.array_equal(rp_o, rp_f) and np.array_equal(c_o, c_f)
    assert rp_f[1] - rp_f[0] == 2 * (n + 1)  # the centre's rows couple it with every rim point
    err = rel_err(v_f, v_o)
    print(f"fan {n} {prec}: |K - K_oracle| / max|K| = {err:.2e}", flush=True)
    assert err < (1e-12 if prec == "f64" else 2e-5)
    f.assemble_kk()
    assert np.array_equal(f.get_csr()[2], v_f)
