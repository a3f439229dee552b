"""Serving: many independent footing solves admitted concurrently on one GPU (pm_serve, pm_server_*).

Each request's result must be bitwise the same whether it runs alone or in the mix (its kernels, their
order and its stream are its own), admission must keep the sum of running footprints within the HBM
budget, a failing request must not disturb the others, and served solutions must match the oracle.
"""
import numpy as np
import pytest

from gpu_util import rel_err
from oracle import oracle as orc
from pmsim_amd import describe
from pmsim_amd.solver import Fem, Server, serve
from pmsim_amd.workloads import footing_request

pytestmark = pytest.mark.gpu


def _requests(n, seed, nx=(8, 48), ny=(4, 24)):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        (mesh, base, cfg, ess, nat), p = footing_request(rng, nx, ny)
        cfg.set_scaled_rr(1e-10).set_pcg(1e-12, 50000)
        out.append((describe(mesh, base, cfg, ess, nat), cfg, p))
    return out


def _handles(reqs, keep_history=False):
    return [Fem(d, cfg, keep_history=keep_history, initialize=False) for d, cfg, _ in reqs]


def test_served_results_same_alone_and_in_the_mix():
    reqs = _requests(10, 7)
    mix = _handles(reqs)
    res = serve(mix, np.zeros(len(mix)), max_active=8)
    assert all(r["status"] == 0 for r in res), [r["error"] for r in res]
    admits = sorted(r["t_admit"] for r in res)
    assert admits[7] < min(r["t_done"] for r in res)  # the first 8 really ran together
    for k, (d, cfg, p) in enumerate(reqs):
        alone = _handles([reqs[k]])
        r = serve(alone, [0.0], max_active=1)[0]
        assert r["status"] == 0
        assert r["n_iterations"] == res[k]["n_iterations"] and r["n_pcg_iterations"] == res[k]["n_pcg_iterations"]
        (ua, la), (um, lm) = alone[0].get_u(), mix[k].get_u()
        assert np.array_equal(ua, um) and np.array_equal(la, lm)
        assert np.array_equal(alone[0].get_ip_state(), mix[k].get_ip_state())


def test_admission_respects_hbm_budget():
    reqs = _requests(12, 11)
    hs = _handles(reqs)
    est = [h.estimate_device_bytes() for h in hs]
    budget = int(2.2 * max(est))
    res = serve(hs, np.linspace(0.0, 0.05, len(hs)), hbm_budget_bytes=budget, max_active=12)
    assert all(r["status"] == 0 for r in res)
    for r in res:  # at every admission, the running footprints (including the admitted one) fit
        t = r["t_admit"]
        running = sum(q["device_bytes"] for q in res if q["t_admit"] <= t < q["t_done"])
        assert running <= budget
        assert r["t_admit"] >= r["t_arrival"] - 1e-6


def test_failing_request_is_isolated():
    reqs = _requests(4, 13)
    hs = _handles(reqs)
    bad = reqs[1][0].copy()
    bad["params"] = {a: (1, E, nu, H, 0.0, ng) for a, (m, E, nu, H, z0, ng) in bad["params"].items()}  # z_ini = 0
    hs[1] = Fem(bad, reqs[1][1], keep_history=False, initialize=False)
    res = serve(hs, np.zeros(4), max_active=4)
    assert res[1]["status"] != 0 and "yield surface" in res[1]["error"]
    assert all(res[k]["status"] == 0 for k in (0, 2, 3))


def test_served_solutions_match_oracle():
    reqs = _requests(3, 17, nx=(8, 24), ny=(4, 12))
    hs = _handles(reqs, keep_history=True)
    res = serve(hs, [0.0, 0.01, 0.02], max_active=3)
    for (d, cfg, p), h, r in zip(reqs, hs, res):
        assert r["status"] == 0
        uo, ipo, _ = orc.Oracle(d).solve()
        hist = h.history()
        assert len(hist) == len(uo)
        assert rel_err(hist[-1][1], uo[-1][: h.ndof]) < 1e-9
        assert rel_err(h.get_u()[0], uo[-1][: h.ndof]) < 1e-9


def test_batched_pcg_matches_standalone_solve():
    """Requests large enough for the multi-workgroup PCG go through the batched PCG engine when served;
    their iterates are the single-request ones, so served results equal pm_solve's bit for bit."""
    reqs = _requests(6, 19, nx=(40, 80), ny=(16, 40))
    hs = _handles(reqs)
    res = serve(hs, np.zeros(len(hs)), max_active=6)
    assert all(r["status"] == 0 for r in res)
    for (d, cfg, p), h, r in zip(reqs, hs, res):
        assert h.ndof > 1536  # past the single-workgroup PCG
        f = Fem(d, cfg, keep_history=False)
        f.solve()
        assert f.stats()["n_pcg_iterations"] == r["n_pcg_iterations"]
        assert np.array_equal(f.get_u()[0], h.get_u()[0])
        assert np.array_equal(f.get_ip_state(), h.get_ip_state())


@pytest.mark.parametrize("name", ["footing_f64", "footing_f32", "footing_bf16", "hex20_f64", "arc_lmm",
                                  "footing_f64_multi", "footing_f32_multi"])
def test_footprint_estimate_bounds_the_allocation(name):
    """Admission relies on pm_estimate_device_bytes being an upper bound of what the handle allocates
    (including the arc-length vectors allocated at the first arc-length solve, and the full-pattern K
    copy allocated at the first multi-workgroup PCG solve: the *_multi cases force that PCG)."""
    import cases
    case = {"footing_f64": cases.footing_c1(n_inc=2), "footing_f32": cases.footing_c1(n_inc=2),
            "footing_bf16": cases.footing_c1(n_inc=2), "hex20_f64": cases.hex20_cube(3, 3, 2, n_inc=2),
            "arc_lmm": cases.footing_arc(lmm=True, n_inc=3), "footing_f64_multi": cases.footing_c1(n_inc=2),
            "footing_f32_multi": cases.footing_c1(n_inc=2)}[name]
    mesh, base, cfg, ess, nat = case
    cfg.set_tangent_precision(name.split("_")[1] if name.startswith("footing") else "f64")
    if name.endswith("_multi"):
        cfg.set_pcg_small_max(0)
    f = Fem(describe(mesh, base, cfg, ess, nat), cfg, keep_history=False, initialize=False)
    est = f.estimate_device_bytes()
    f.initialize()
    f.solve()
    used = f.stats()["device_bytes"]
    assert 0 < used <= est, (used, est)
    if name.endswith("_multi"):
        assert f.stats()["n_pcg_iterations"] > 0


def test_batched_pcg_matches_standalone_solve_hex20():
    """The same on Hex20 meshes, where the single-request SpMV takes two rows per lane-group pass and the
    batched one a row at a time: each row's sums and the p.q order are the same, so are the bits."""
    import cases
    hs, built = [], []
    # 20x20x20 has 35,721 nodes: more block rows than the SpMV's 32,768 lane groups, so its groups pair rows
    for k, (nx, ny, nz) in enumerate([(4, 4, 3), (20, 20, 20)]):
        # 0.1 % compression, below first yield: a few Newton iterations (the bits, not plasticity, are tested)
        mesh, base, cfg, ess, nat = cases.hex20_cube(nx, ny, nz, n_inc=2, seed=100 + k, uz=-0.001)
        cfg.set_pcg_small_max(0)  # every system through the multi-workgroup PCG
        d = describe(mesh, base, cfg, ess, nat)
        built.append((d, cfg))
        hs.append(Fem(d, cfg, keep_history=False, initialize=False))
    res = serve(hs, np.zeros(len(hs)), max_active=2)
    assert all(r["status"] == 0 for r in res), [r.get("error") for r in res]
    assert hs[1].ndof == 3 * 35721
    for (d, cfg), h, r in zip(built, hs, res):
        f = Fem(d, cfg, keep_history=False)
        f.solve()
        assert r["n_pcg_iterations"] > 0
        assert f.stats()["n_pcg_iterations"] == r["n_pcg_iterations"]
        assert np.array_equal(f.get_u()[0], h.get_u()[0])
        assert np.array_equal(f.get_ip_state(), h.get_ip_state())


def test_served_iteration_limit_matches_standalone():
    """A PCG limit that is not a multiple of the engine's 16-iteration chunk (40, reached before convergence):
    the batched engine stops each system at exactly 40 iterations as the single-request PCG does, both fall
    back to the host direct solver, and served results equal standalone ones bit for bit."""
    reqs = _requests(3, 23, nx=(40, 60), ny=(16, 24))
    for d, cfg, p in reqs:
        cfg.set_pcg(1e-12, 40).set_pcg_small_max(0)
    hs = _handles(reqs)
    res = serve(hs, np.zeros(len(hs)), max_active=3)
    assert all(r["status"] == 0 for r in res), [r["error"] for r in res]
    for (d, cfg, p), h, r in zip(reqs, hs, res):
        f = Fem(d, cfg, keep_history=False)
        f.solve()
        st = f.stats()
        assert st["n_host_fallbacks"] > 0
        assert st["n_pcg_iterations"] == r["n_pcg_iterations"]
        assert np.array_equal(f.get_u()[0], h.get_u()[0])


def test_served_fp32_tangent_within_bar_of_fp64():
    """The mixed-precision served form (FP32 element tangents, K held in FP32; residual, return map and PCG in
    FP64) against the FP64 served form on the same requests, one of them at the top of the size range
    (256 x 128 Qua8 = 32,768 elements, 198k DOFs): every final displacement within the 1e-8 relative bar."""
    from pmsim_amd.workloads import strip_footing
    reqs = _requests(4, 29)
    mesh, base, cfg, ess, nat = strip_footing(256, 128, n_inc=5, young=2e4, poisson=0.3, z_ini=150.0, width=1.5)
    cfg.set_scaled_rr(1e-10).set_pcg(1e-12, 50000)
    reqs.append((describe(mesh, base, cfg, ess, nat), cfg, None))
    u = {}
    for prec in ("f64", "f32"):
        for d, c, p in reqs:
            c.set_tangent_precision(prec)
        hs = _handles(reqs)
        res = serve(hs, np.zeros(len(hs)), max_active=5)
        assert all(r["status"] == 0 for r in res), [r["error"] for r in res]
        u[prec] = [h.get_u()[0] for h in hs]
    assert hs[-1].ndof == 2 * (2 * 256 + 1) * (2 * 128 + 1) - 2 * 256 * 128
    errs = [rel_err(b, a) for a, b in zip(u["f64"], u["f32"])]
    print("served FP32-tangent vs FP64 u:", " ".join(f"{e:.2e}" for e in errs), flush=True)
    assert max(errs) < 1e-8


def test_open_submission_batches_passes_and_matches_standalone():
    """pm_server_*: requests submitted while earlier ones run (the first is not done right after its submission),
    the element passes of running requests share launches (fewer batches than passes), and every result is
    bitwise the one of its own pm_solve (the shared launches do the same per-element arithmetic)."""
    import time
    reqs = _requests(8, 31, nx=(16, 60), ny=(8, 30))
    hs = _handles(reqs)
    with Server(max_active=8) as srv:
        tickets = [srv.submit(h) for h in hs[:4]]
        assert srv.poll(tickets[0]) is None
        time.sleep(0.05)
        tickets += [srv.submit(h) for h in hs[4:]]  # arrive while the first ones run
        res = [srv.wait(t) for t in tickets]
        st = srv.stats()
    assert all(r["status"] == 0 for r in res), [r["error"] for r in res]
    assert all(r["t_admit"] >= r["t_arrival"] and r["t_done"] > r["t_admit"] for r in res)
    assert st["pass_jobs"] > st["pass_batches"] > 0, st
    for (d, cfg, p), h, r in zip(reqs, hs, res):
        f = Fem(d, cfg, keep_history=False)
        f.solve()
        assert f.stats()["n_iterations"] == r["n_iterations"]
        assert np.array_equal(f.get_u()[0], h.get_u()[0])
        assert np.array_equal(f.get_ip_state(), h.get_ip_state())


def test_server_refuses_initialized_handles_and_unknown_tickets():
    reqs = _requests(1, 37)
    f = Fem(reqs[0][0], reqs[0][1], keep_history=False)  # initialized
    with Server() as srv:
        with pytest.raises(Exception, match="not yet initialized"):
            srv.submit(f)
        with pytest.raises(Exception, match="unknown ticket"):
            srv.poll(5)
