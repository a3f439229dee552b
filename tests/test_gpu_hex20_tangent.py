"""The persistent Hex20 tangent (27 IPs) against the node-block vector kernel with one workgroup per element
group: the same closed form (the persistent kernel sums its elastic IPs as one Gram turned into their terms
once), so the element matrices, the assembled K and K v agree to rounding (1e-14 relative to the largest
entry), in FP64 and FP32 storage, on meshes whose element counts leave the persistent
wavefronts with zero, one, a few and about forty elements each (the next element's data is prefetched into
LDS while the current one is accumulated, so more than one element per wavefront exercises the prefetch; a
stale or half-written prefetch would show as an O(1) difference).  The persistent kernel is deterministic
(two passes bitwise equal) and matches the oracle at 1e-12.  The persistent matrix-core kernels (FP32 and BF16
operands) are checked against the FP64 K and against the per-group matrix-core kernel."""
import numpy as np
import pytest

from gpu_util import rel_err
from oracle import oracle as orc
from pmsim_amd import Config, Essential, FemBase, Natural, ParamSolid, StressStrain, describe
from pmsim_amd import generators as G
from pmsim_amd.solver import Fem

pytestmark = pytest.mark.gpu
TOL = {"f64": 1e-14, "f32": 1e-7}  # FP32: one rounding of the stored values may differ by an ulp


def _pair_pass(n, prec, seed=7):
    """Both kernels after the same partly plastic increment of a jittered Hex20 block n = (nx, ny, nz)."""
    mesh = G.hex20_block(*n, 1.0, 1.0, 1.0, seed=20261203, jitter=0.05)
    base = FemBase(mesh, [(1, ParamSolid(1.0, StressStrain.VonMises(2e5, 0.3, 1e3, 250.0), 27))])
    X = mesh.coords
    rng = np.random.default_rng(seed)
    u0 = np.zeros_like(X)
    u0[:, 2] = -2.5e-3 * X[:, 0] * X[:, 2]
    u0 = u0.ravel() + rng.uniform(-2e-6, 2e-6, X.size)
    out = []
    for kern in ("vector", "vector-wg"):
        cfg = Config(mesh).set_tangent_precision(prec).set_tangent_kernel(kern)
        f = Fem(describe(mesh, base, cfg, Essential(), Natural()), cfg, keep_history=False)
        f.backup_secondary_values()
        f.set_ddu(u0)
        f.element_pass()
        out.append(f)
    return mesh, u0, out


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("n", [(3, 2, 2), (5, 7, 3), (20, 20, 10)])
def test_persistent_hex20_tangent_vs_per_group(n, prec):
    mesh, u0, (fp, fw) = _pair_pass(n, prec)
    plastic = np.mean(fp.get_ip_state()[:, fp.ns + 2] == 0.0)
    assert 0.0 < plastic < 1.0
    rp, cp, vp = fp.get_csr()
    rw, cw, vw = fw.get_csr()
    assert np.array_equal(rp, rw) and np.array_equal(cp, cw)
    err = rel_err(vp, vw)
    print(f"{n} {prec}: |K (persistent) - K (per-workgroup)| / max |K| = {err:.2e}", flush=True)
    assert err < TOL[prec]
    for cell in sorted({0, 1, mesh.ncell // 2, mesh.ncell - 1}):
        assert rel_err(fp.element_kk(cell), fw.element_kk(cell)) < TOL[prec]
    fp.element_pass()  # repeatable
    assert np.array_equal(fp.get_csr()[2], vp)


def test_persistent_hex20_tangent_matches_oracle():
    """Element matrices of the persistent kernel against the oracle's dense B^T (D B) after a partly plastic
    update (stored in the cell's node order, read back through the ABI)."""
    mesh, u0, (fp, fw) = _pair_pass((3, 2, 2), "f64")
    base = FemBase(mesh, [(1, ParamSolid(1.0, StressStrain.VonMises(2e5, 0.3, 1e3, 250.0), 27))])
    d = describe(mesh, base, Config(mesh), Essential(), Natural())
    o = orc.Oracle(d)
    o.init_state()
    o.update_secondary(u0)  # from the initial state, as the handles' update from their backup
    rp_o, c_o, v_o = o.kk_csr()
    rp_f, c_f, v_f = fp.get_csr()
    assert np.array_equal(rp_o, rp_f) and np.array_equal(c_o, c_f)
    assert rel_err(v_f, v_o) < 1e-12


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_persistent_hex20_tangent_full_size(prec):
    """64x64x32 Hex20 (131,072 elements: about 43 per wavefront): K v of the two kernels equal to rounding for
    a random v, and the element matrices of sampled cells likewise (cells at the ends of wavefront ranges
    included)."""
    mesh, u0, (fp, fw) = _pair_pass((64, 64, 32), prec)
    v = np.random.default_rng(3).normal(size=fp.ndof)
    kp, kw = fp.apply_kk(v), fw.apply_kk(v)
    print(f"64x64x32 {prec}: max |K v (persistent) - K v (per-workgroup)| = {np.abs(kp - kw).max():.3e}", flush=True)
    assert rel_err(kp, kw) < 10 * TOL[prec]
    rng = np.random.default_rng(5)
    for cell in [0, 42, 43, 44, 85, 86, mesh.ncell - 1] + list(rng.integers(0, mesh.ncell, 8)):
        assert rel_err(fp.element_kk(int(cell)), fw.element_kk(int(cell))) < TOL[prec], f"cell {cell}"
    fp.element_pass()
    assert np.array_equal(fp.apply_kk(v), kp)


def _mfma_pair(n, prec, seed=7):
    """FP64 handle and the two reduced-precision matrix-core kernels (persistent Hex20 form, per-group form)
    after the same partly plastic increment."""
    mesh = G.hex20_block(*n, 1.0, 1.0, 1.0, seed=20261203, jitter=0.05)
    base = FemBase(mesh, [(1, ParamSolid(1.0, StressStrain.VonMises(2e5, 0.3, 1e3, 250.0), 27))])
    X = mesh.coords
    rng = np.random.default_rng(seed)
    u0 = np.zeros_like(X)
    u0[:, 2] = -2.5e-3 * X[:, 0] * X[:, 2]
    u0 = u0.ravel() + rng.uniform(-2e-6, 2e-6, X.size)
    out = []
    for p, kern in (("f64", "default"), (prec, "mfma"), (prec, "mfma-wg")):
        cfg = Config(mesh).set_tangent_precision(p).set_tangent_kernel(kern)
        f = Fem(describe(mesh, base, cfg, Essential(), Natural()), cfg, keep_history=False)
        f.backup_secondary_values()
        f.set_ddu(u0)
        f.element_pass()
        out.append(f)
    return mesh, out


# bars relative to the largest |K|: the matrix-core kernels' bars (FP32 operands and accumulation 2e-5, BF16
# operands with FP32 accumulation 3e-2), and the slack over the per-group kernel's own distance from the FP64 K
MFMA_TOL = {"f32": (2e-5, 5e-6), "bf16": (3e-2, 1e-2)}
KV_TOL = {"f32": 1e-5, "bf16": 1e-2}


@pytest.mark.parametrize("prec", ["f32", "bf16"])
@pytest.mark.parametrize("n", [(3, 2, 2), (5, 7, 3), (20, 20, 10)])
def test_persistent_hex20_mfma_tangent(n, prec):
    """FP32 or BF16 operands with FP32 accumulation: K of the persistent kernel within the matrix-core bar of
    the FP64 K and no further from it than the per-group kernel's K plus the operand rounding of either;
    diagonal node blocks exactly symmetric; f_int bitwise the FP64 handle's (the stress update and B^T sigma
    stay FP64); repeatable."""
    mesh, (f64, fp, fw) = _mfma_pair(n, prec)
    v64 = f64.get_csr()[2]
    vp, vw = fp.get_csr()[2], fw.get_csr()[2]
    ep, ew = rel_err(vp, v64), rel_err(vw, v64)
    print(f"{n} {prec}: |K - K64| / max|K64|: persistent {ep:.2e}, per-group {ew:.2e}", flush=True)
    bar, slack = MFMA_TOL[prec]
    assert ep < bar and ep < ew + slack
    assert np.array_equal(fp.get_f_int(), f64.get_f_int())
    for cell in sorted({0, mesh.ncell // 2, mesh.ncell - 1}):
        k = fp.element_kk(cell)
        for a in range(20):
            blk = k[3 * a:3 * a + 3, 3 * a:3 * a + 3]
            assert np.array_equal(blk, blk.T)
    fp.element_pass()
    assert np.array_equal(fp.get_csr()[2], vp)


@pytest.mark.parametrize("prec", ["f32", "bf16"])
def test_persistent_hex20_mfma_tangent_full_size(prec):
    """64x64x32 (131,072 elements): K v of the persistent matrix-core kernel within the bar of the FP64 K v
    (relative, max norm: 1e-5 FP32, 1e-2 BF16) and repeatable bit for bit."""
    mesh, (f64, fp, fw) = _mfma_pair((64, 64, 32), prec)
    v = np.random.default_rng(3).normal(size=f64.ndof)
    k64, kp, kw = f64.apply_kk(v), fp.apply_kk(v), fw.apply_kk(v)
    ep, ew = rel_err(kp, k64), rel_err(kw, k64)
    print(f"64x64x32 {prec}: |K v - K64 v| / max|K64 v|: persistent {ep:.2e}, per-group {ew:.2e}", flush=True)
    assert ep < KV_TOL[prec]
    fp.element_pass()
    assert np.array_equal(fp.apply_kk(v), kp)
