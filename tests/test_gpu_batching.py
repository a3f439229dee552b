"""GPU parity of the batched path on uneven element mixes, and of captured-graph replays.

The mixed 2D mesh (30% Qua8, 25% Qua4, 20% Tri6, 15% Qua9, 10% Tri15) runs as five per-kind batches;
the oracle evaluates it element by element (per-element scalar dispatch).  Stresses agree to 1e-12
relative (to the largest |sigma|) on every kind but Tri15, and to 5e-11 on Tri15: its strain
Sum_m u_m G_m cancels terms about 10^3-10^4 times larger than the result when the increment is a
large-magnitude linear field (u ~ 1 over a 100 m domain, strain ~ 1e-2), so two correct FP64 summation
orders differ at that level (the shape functions and Gauss rules of both sides agree to 1e-15, the
element Jacobians have condition numbers below 4).  f_int and K mix the kinds: 5e-11.  The CSR pattern
is bit-exact.  Graph replays must reproduce eager launches bit for bit, replay after replay, with the
program's own eager work in between.
"""
import numpy as np
import pytest

from gpu_util import rel_err
from oracle import oracle as orc
from pmsim_amd import Config, describe
from pmsim_amd import mesh as M
from pmsim_amd.solver import Fem
from pmsim_amd.workloads import mixed_compression, strip_footing

pytestmark = pytest.mark.gpu


def _ddu(mesh, scale, seed):
    rng = np.random.default_rng(seed)
    A = rng.uniform(-scale, scale, (2, 2))
    return ((mesh.coords @ A.T) + rng.uniform(-scale / 5, scale / 5, mesh.coords.shape)).ravel()


@pytest.mark.parametrize("kernel", ["default", "mfma"])
def test_mixed_200k_element_pass_vs_oracle(kernel):
    """Full-size mixed mesh (200,165 elements, 1.40 M DOF): return map, f_int and K of one pass from a
    partly plastic increment against the oracle's per-element evaluation."""
    mesh, base, cfg, ess, nat = mixed_compression(200_000)
    cfg.set_tangent_kernel(kernel)
    d = describe(mesh, base, cfg, ess, nat)
    o = orc.Oracle(d)
    f = Fem(d, cfg, keep_history=False)
    o.init_state()
    ddu = _ddu(mesh, 1.2e-2, 5)
    o.update_secondary(ddu)
    f.set_ddu(ddu)
    f.element_pass()
    so, sf = o.get_ip_state(), f.get_ip_state()
    plastic = so[:, o.ns + 2] == 0.0
    assert 0.2 < plastic.mean() < 0.95
    assert np.array_equal(plastic, sf[:, o.ns + 2] == 0.0)
    nip_of = np.array([len(orc.gauss(int(k))) for k in mesh.kinds])
    tri15 = np.repeat(mesh.kinds == M.TRI15, nip_of)
    smax = np.abs(so[:, : o.ns]).max()
    dev = np.abs(sf[:, : o.ns] - so[:, : o.ns]).max(axis=1) / smax
    assert dev[~tri15].max() < 1e-12
    assert dev[tri15].max() < 5e-11
    assert rel_err(f.get_f_int(), o.f_int()) < 5e-11
    rp_o, c_o, v_o = o.kk_csr()
    rp_f, c_f, v_f = f.get_csr()
    assert np.array_equal(rp_o, rp_f) and np.array_equal(c_o, c_f)
    assert rel_err(v_f, v_o) < 5e-11


def test_mixed_newton_vs_oracle():
    """Mixed mesh (~800 elements) compressed far enough to yield: Newton with PCG on the GPU against the
    oracle's literal LMM Newton with a direct solver, every increment."""
    mesh, base, cfg, ess, nat = mixed_compression(800, n_inc=4, uy=-1.5)
    cfg.set_pcg(1e-13, 50000)
    d = describe(mesh, base, cfg, ess, nat)
    f = Fem(d, cfg)
    f.solve()
    hist = f.history()
    uo, ipo, _ = orc.Oracle(d).solve()
    assert len(hist) == len(uo)
    for k in range(len(hist)):
        assert rel_err(hist[k][1], uo[k][: f.ndof]) < 1e-9
    ns = f.ns
    assert (ipo[-1][:, ns + 2] == 0.0).mean() > 0.05  # yielded
    assert rel_err(hist[-1][3][:, :ns], ipo[-1][:, :ns]) < 1e-8


def test_graph_replays_match_eager_element_pass():
    mesh, base, cfg, ess, nat = mixed_compression(20_000)
    d = describe(mesh, base, cfg, ess, nat)
    ddu = _ddu(mesh, 1.2e-2, 9)
    res = {}
    for graphs in (False, True):
        c = Config(mesh).set_use_graphs(graphs)
        f = Fem(d, c, keep_history=False)
        f.set_ddu(ddu)
        outs = []
        for rep in range(3):  # replays, with eager host work and synchronisation in between
            f.element_pass()
            outs.append((f.get_ip_state(), f.get_f_int(), f.get_csr()[2]))
            f.set_ddu(ddu * (1.0 + 0.0 * rep))
        res[graphs] = outs
    for rep in range(3):
        for a, b in zip(res[False][rep], res[True][rep]):
            assert np.array_equal(a, b)


def test_graph_pcg_newton_matches_eager():
    """Strip footing 96x48 Qua8 (29 k DOF, the multi-workgroup PCG path): the Newton history with
    captured 16-iteration PCG chunks and element passes is bitwise the eager one."""
    hists = {}
    for graphs in (False, True):
        mesh, base, cfg, ess, nat = strip_footing(96, 48, n_inc=4)
        cfg.set_use_graphs(graphs).set_pcg(1e-12, 50000)
        d = describe(mesh, base, cfg, ess, nat)
        f = Fem(d, cfg)
        f.solve()
        hists[graphs] = f.history()
        assert f.stats()["n_pcg_iterations"] > 100
    assert len(hists[False]) == len(hists[True])
    for a, b in zip(hists[False], hists[True]):
        assert np.array_equal(a[1], b[1]) and np.array_equal(a[3], b[3])


def test_tuner_on_mixed_batches():
    """The tuner picks a kernel per element batch; the tuned K equals the default one to rounding."""
    mesh, base, cfg, ess, nat = mixed_compression(20_000)
    d = describe(mesh, base, cfg, ess, nat)
    ddu = _ddu(mesh, 1.2e-2, 11)
    vals = []
    for k in ("default", "tune"):
        f = Fem(d, Config(mesh).set_tangent_kernel(k), keep_history=False)
        f.set_ddu(ddu)
        f.element_pass()
        vals.append(f.get_csr()[2])
    assert rel_err(vals[1], vals[0]) < 1e-13
