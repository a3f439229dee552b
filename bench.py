"""Benchmark: Hex20 IP-updates/s of the implicit element pass (K + R + return map) on one MI355X.

Workload: unit cube, structured 64x64x32 Hex20 (131,072 elements, 27 IPs each = 3,538,944 IPs,
549,185 points, 1,647,555 DOFs), corner jitter ±5% of h (seed 20261203), von Mises E=2e5 nu=0.3
z0=250 H=1e3.  A step is one element pass over device-resident data: gather of coordinates and
displacement increments, strain increment, radial return from the backup state (with per-IP
sigma, z, lambda, loading flag, epbar), B^T sigma into the element forces and their deterministic
gather into R, B^T D B node blocks and their deterministic gather into the node-block CSR tangent.
The increment is a seeded field large enough that part of the IPs return plastically.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--nx 64 --ny 64 --nz 32] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
PEAK_TFLOPS = {"f64": 78.6, "f32": 157.3, "bf16": 2500.0}  # MI355X dense FP64 / FP32 / BF16 matrix peaks
FLOP_PER_IP = 2 * 60 * 60 * 6 + 2 * 6 * 6 * 60  # dense B^T D B per Hex20 IP = 47,520
# what the closed-form tangent executes per Hex20 IP: 210 upper node blocks x 30 FMA (three rank-1
# updates and the diagonal dot product), the per-tile scalings (55 tiles x 18) and the per-IP set-up
# (J, J^-1, D coefficients, node vectors: about 1,200)
EXECUTED_FLOP_PER_IP = 210 * 30 * 2 + 55 * 18 + 1200


def build_problem(nx, ny, nz, precision="f64", kernel="default"):
    from pmsim_amd import Config, Essential, FemBase, Natural, ParamSolid, StressStrain, describe
    from pmsim_amd import generators as G
    mesh = G.hex20_block(nx, ny, nz, 1.0, 1.0, 1.0, seed=20261203, jitter=0.05)
    base = FemBase(mesh, [(1, ParamSolid(1.0, StressStrain.VonMises(2e5, 0.3, 1e3, 250.0), 27))])
    cfg = Config(mesh).set_tangent_precision(precision).set_tangent_kernel(kernel)
    return mesh, base, cfg, describe(mesh, base, cfg, Essential(), Natural())


def increment(mesh, seed=7):
    """Vertical compression growing with x (0 at x=0 to 0.25% at x=1) plus seeded noise: the IPs on the
    high-x side return plastically, the others stay elastic, so both branches of the return map run."""
    rng = np.random.default_rng(seed)
    X = mesh.coords
    u = np.zeros_like(X)
    u[:, 2] = -2.5e-3 * X[:, 0] * X[:, 2]
    u += rng.uniform(-2e-6, 2e-6, X.shape)
    return u.ravel()


def measured_traffic(nx, ny, nz, precision, kernel):
    """HBM bytes per launch of the tangent kernel from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE
    summary of this workload (scaled by the copy-kernel calibration stored with it), or None."""
    if (nx, ny, nz) != (64, 64, 32):
        return None, None
    tag = {("f64", "default"): "f64_vector", ("f64", "vector"): "f64_vector", ("f64", "mfma"): "f64_mfma",
           ("bf16", "default"): "bf16_mfma", ("bf16", "mfma"): "bf16_mfma"}.get((precision, kernel))
    path = os.path.join(ROOT, "profiles", f"element_pass_hex20_64x64x32_{tag}.json")
    if tag is None or not os.path.exists(path):
        return None, None
    prof = json.load(open(path))
    for name, k in prof["kernels"].items():
        if name.startswith("k_tangent"):
            return k["read_bytes"] + k["write_bytes"], name
    return None, None


def _host_cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _host_threads():
    """Cores this process may use: the affinity mask, capped by OMP_NUM_THREADS when that is set."""
    n = len(os.sched_getaffinity(0))
    cap = os.environ.get("OMP_NUM_THREADS", "")
    return max(1, min(n, int(cap))) if cap.isdigit() and int(cap) > 0 else n


def cpu_baseline(seconds=12.0, seconds_all=6.0):
    """Oracle element pass (update + f_int + K per element) over a bounded sample: on one host core (the
    reported value) and, as `all_cores`, with its OpenMP element loops on every core this process may use."""
    from oracle import oracle as orc
    mesh, base, cfg, d = build_problem(8, 8, 4)
    o = orc.Oracle(d)
    o.init_state()
    ddu = increment(mesh)
    t1 = o.time_element_pass(ddu, reps=1, nthreads=1)
    reps = max(1, int(seconds / max(t1, 1e-6)))
    t = o.time_element_pass(ddu, reps=reps, nthreads=1)
    nip = mesh.ncell * 27
    out = dict(value=nip * reps / t, unit="IP-updates/s", cores=1, kind="port",
               sample=f"8x8x4 Hex20 (256 elements, {nip} IPs) x {reps} passes, serial C++ oracle (update+f_int+K)",
               host=_host_cpu_model())
    nt = _host_threads()
    if nt > 1:
        mesh2, _, _, d2 = build_problem(16, 16, 8)
        o2 = orc.Oracle(d2)
        o2.init_state()
        ddu2 = increment(mesh2)
        t1 = o2.time_element_pass(ddu2, reps=1, nthreads=nt)
        reps2 = max(1, int(seconds_all / max(t1, 1e-6)))
        t2 = o2.time_element_pass(ddu2, reps=reps2, nthreads=nt)
        nip2 = mesh2.ncell * 27
        out["all_cores"] = dict(value=nip2 * reps2 / t2, unit="IP-updates/s", cores=nt,
                                sample=f"16x16x8 Hex20 ({mesh2.ncell} elements, {nip2} IPs) x {reps2} passes, "
                                       f"C++ oracle with OpenMP element loops on {nt} threads")
    return out


def run_mixed2d(a):
    """Mixed 2D mesh (200k elements in five per-kind batches): element pass replayed as a captured graph,
    and optionally the full incremental Newton solve (8 increments, PCG)."""
    from pmsim_amd import describe
    from pmsim_amd.solver import Fem
    from pmsim_amd.workloads import mixed_compression
    t0 = time.time()
    mesh, base, cfg, ess, nat = mixed_compression(a.elements)
    cfg.set_tangent_precision(a.precision).set_tangent_kernel(a.kernel).set_scaled_rr(1e-10)
    d = describe(mesh, base, cfg, ess, nat)
    f = Fem(d, cfg, keep_history=False)
    kinds, counts = np.unique(mesh.kinds, return_counts=True)
    print(f"setup: {mesh.ncell} cells, {f.nip} IPs, {f.ndof} DOFs, {time.time() - t0:.1f} s", file=sys.stderr, flush=True)
    rng = np.random.default_rng(5)
    A = rng.uniform(-1.2e-2, 1.2e-2, (2, 2))
    f.backup_secondary_values()
    f.set_ddu(((mesh.coords @ A.T) + rng.uniform(-2.4e-3, 2.4e-3, mesh.coords.shape)).ravel())
    for _ in range(a.warmup):
        f.element_pass()
    ms = f.time_element_pass(a.steps)
    plastic = float(np.mean(f.get_ip_state()[:, f.ns + 2] == 0.0))
    out = {"metric": "mixed 2D IP-updates/s (K+R+return map)", "value": f.nip / (ms["pass_"] * 1e-3), "unit": "IP-updates/s",
           "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": ms["pass_"], "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64" if a.precision == "f64" else f"{a.precision}-tangent/f64",
           "data": "synthetic (seeded mixed 2D mesh, seeded increment)",
           "config": {"workload": f"mixed 2D mesh {mesh.ncell} elements (Qua8/Qua4/Tri6/Qua9/Tri15), one element pass",
                      "elements": int(mesh.ncell), "ips": int(f.nip), "dofs": int(f.ndof), "plastic_fraction": plastic,
                      "kinds": {M_NAMES[int(k)]: int(c) for k, c in zip(kinds, counts)}},
           "kernels_ms": {k: ms[k] for k in ("update", "residual", "tangent", "scatter", "total")}}
    if a.solve:
        g = Fem(d, cfg, keep_history=False)
        t1 = time.time()
        g.solve()
        st = g.stats()
        out["solve"] = {"increments": 8, "seconds": time.time() - t1, "newton_iterations": st["n_iterations"],
                        "pcg_iterations": st["n_pcg_iterations"], "ms_linear": st["ms_linear"], "ms_total": st["ms_total"]}
    print(json.dumps(out), flush=True)


def run_serve(a):
    """Serving: Poisson arrivals of parametric strip-footing solves, admitted by HBM footprint and solved
    concurrently (pm_serve); solves/s and latency percentiles, FP64 and FP32-tangent runs of the same
    trace, and the largest relative difference of their final displacements."""
    from pmsim_amd import describe
    from pmsim_amd.solver import Fem, serve
    from pmsim_amd.workloads import footing_request, poisson_arrivals
    rng = np.random.default_rng(a.seed)
    reqs = []
    t0 = time.time()
    for _ in range(a.requests):
        (mesh, base, cfg, ess, nat), p = footing_request(rng)
        reqs.append((mesh, base, cfg, ess, nat, p))
    arrivals = poisson_arrivals(np.random.default_rng(a.seed + 1), a.requests, a.rate)
    print(f"setup: {a.requests} requests, {sum(r[0].ncell for r in reqs)} elements in all, {time.time() - t0:.1f} s",
          file=sys.stderr, flush=True)
    # untimed warm-up: the first launches of each kernel load its code object
    wrng = np.random.default_rng(1)
    for prec in ("f64", "f32"):
        (mesh, base, cfg, ess, nat), _ = footing_request(wrng, (8, 12), (4, 6))
        cfg.set_tangent_precision(prec)
        serve([Fem(describe(mesh, base, cfg, ess, nat), cfg, keep_history=False, initialize=False)], [0.0])
    runs = {}
    for prec in ("f64", "f32"):
        hs = []
        for mesh, base, cfg, ess, nat, p in reqs:
            cfg.set_scaled_rr(1e-10).set_pcg(1e-12, 50000).set_tangent_precision(prec)
            hs.append(Fem(describe(mesh, base, cfg, ess, nat), cfg, keep_history=False, initialize=False))
        res = serve(hs, arrivals, hbm_budget_bytes=int(a.hbm_gb * (1 << 30)), max_active=a.max_active)
        ok = [r for r in res if r["status"] == 0]
        lat = np.array([r["t_done"] - r["t_arrival"] for r in ok])
        span = max(r["t_done"] for r in res) - min(arrivals)
        runs[prec] = dict(res=res, u=[h.get_u()[0] if r["status"] == 0 else None for h, r in zip(hs, res)],
                          solves_per_s=len(ok) / span, p50=float(np.percentile(lat, 50)), p99=float(np.percentile(lat, 99)),
                          failed=len(res) - len(ok), newton=sum(r["n_iterations"] for r in ok),
                          pcg=sum(r["n_pcg_iterations"] for r in ok))
        print(f"{prec}: {len(ok)}/{len(res)} solved in {span:.1f} s, p50 {runs[prec]['p50']:.2f} s, p99 {runs[prec]['p99']:.2f} s",
              file=sys.stderr, flush=True)
    du = [float(np.abs(u2 - u1).max() / max(np.abs(u1).max(), 1e-300))
          for u1, u2 in zip(runs["f64"]["u"], runs["f32"]["u"]) if u1 is not None and u2 is not None]
    plain = None
    if a.plain:
        # the plain FP64 GPU form: the same requests in arrival order, one at a time on one stream
        # (set-up + Newton + single-request PCG each), no concurrency and no batching; latencies follow
        # from FIFO service of the same arrival times
        order = np.argsort(arrivals, kind="stable")
        t_free, lat, solve_s, up = 0.0, [], [], []
        for i in order:
            mesh, base, cfg, ess, nat, p = reqs[i]
            cfg.set_scaled_rr(1e-10).set_pcg(1e-12, 50000).set_tangent_precision("f64")
            t1 = time.time()
            h = Fem(describe(mesh, base, cfg, ess, nat), cfg, keep_history=False)
            h.solve()
            u = h.get_u()[0]
            del h
            s = time.time() - t1
            start = max(t_free, float(arrivals[i]))
            t_free = start + s
            lat.append(t_free - float(arrivals[i]))
            solve_s.append(s)
            u64 = runs["f64"]["u"][i]
            if u64 is not None:
                up.append(float(np.abs(u - u64).max() / max(np.abs(u).max(), 1e-300)))
            print(f"plain {len(solve_s)}/{a.requests}: {mesh.ncell} cells, {s:.2f} s", file=sys.stderr, flush=True)
        span = t_free - float(np.min(arrivals))
        plain = {"solves_per_s": a.requests / span, "p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                 "sum_solve_s": float(np.sum(solve_s)), "u_rel_diff_served_vs_alone_max": max(up) if up else None}
    f64 = runs["f64"]
    out = {"metric": "served solves/s", "value": f64["solves_per_s"], "unit": "solves/s", "n_gpus": 1, "steps": a.requests,
           "warmup": 0, "ms_per_step": 1e3 / f64["solves_per_s"], "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded parametric strip footings, Poisson arrivals)",
           "config": {"workload": f"{a.requests} strip-footing solves (Qua8 nx in [8,256], ny in [4,128], 5-40 increments), "
                                  f"Poisson arrivals at {a.rate}/s", "max_active": a.max_active, "hbm_budget_gb": a.hbm_gb,
                      "elements": [int(r[0].ncell) for r in reqs]},
           "latency_s": {"p50": f64["p50"], "p99": f64["p99"]},
           "fp64": {k: f64[k] for k in ("solves_per_s", "p50", "p99", "failed", "newton", "pcg")},
           "fp32_tangent": {k: runs["f32"][k] for k in ("solves_per_s", "p50", "p99", "failed", "newton", "pcg")},
           "u_rel_diff_fp32_vs_fp64": {"max": max(du) if du else None, "median": float(np.median(du)) if du else None}}
    if plain is not None:
        out["plain_fp64_sequential"] = plain
        out["speedup_vs_plain"] = {"served_fp64": f64["solves_per_s"] / plain["solves_per_s"],
                                   "served_fp32_tangent": runs["f32"]["solves_per_s"] / plain["solves_per_s"]}
    print(json.dumps(out), flush=True)


def run_precision(a):
    """Inexact-Newton tangents on Hex20 96x96x48 (5.48 M DOF, 11.9 M IPs): the same incremental solve with
    FP64, FP32 and BF16->FP32-accumulate tangents (residual, return map and PCG vectors stay FP64);
    errors of u, sigma and epbar against the FP64 run, Newton and PCG iteration counts."""
    from pmsim_amd import describe
    from pmsim_amd.solver import Fem
    from pmsim_amd.workloads import hex20_cube
    t0 = time.time()
    if a.uz_stations:
        stations = [0.0] + [float(v) for v in a.uz_stations.split(",")]
        mesh, base, cfg, ess, nat = hex20_cube(a.nx, a.ny, a.nz, stations=stations)
        loading = "top uz at the stations " + ", ".join(f"{v:g}" for v in stations)
    else:
        mesh, base, cfg, ess, nat = hex20_cube(a.nx, a.ny, a.nz, n_inc=a.n_inc, uz=-0.001 * a.n_inc)
        loading = f"{a.n_inc} increments of uz = -0.001"
    print(f"mesh: {mesh.ncell} cells, {mesh.npoint} points, {time.time() - t0:.1f} s", file=sys.stderr, flush=True)
    ref = None
    out = {"metric": "Hex20 inexact-Newton tangent precision", "n_gpus": 1, "data": "synthetic (jittered Hex20 cube, seeded)",
           "config": {"workload": f"Hex20 unit cube {a.nx}x{a.ny}x{a.nz}, 27 IP, von Mises, {loading}",
                      "elements": int(mesh.ncell)}, "runs": {}}
    for prec in a.precisions.split(","):
        cfg.set_tangent_precision(prec).set_scaled_rr(1e-9).set_pcg(1e-9, a.pcg_max).set_n_max_iterations(30).set_verbose(True, True)
        d = describe(mesh, base, cfg, ess, nat)
        f = Fem(d, cfg, keep_history=False)
        print(f"{prec}: handle ready ({f.ndof} DOFs, {f.nnzb} K blocks), {time.time() - t0:.1f} s", file=sys.stderr, flush=True)
        t1 = time.time()
        from pmsim_amd.solver import PmError
        try:
            f.solve()
            failure = None
        except PmError as e:  # reported, not fatal: a reduced-precision tangent may stall Newton or PCG
            failure = str(e)
        secs = time.time() - t1
        st = f.stats()
        if failure is not None:
            out["runs"][prec] = {"failed": failure, "seconds": secs, "converged_steps": st["n_steps"],
                                 "newton_iterations": st["n_iterations"], "pcg_iterations": st["n_pcg_iterations"]}
            print(f"{prec}: {json.dumps(out['runs'][prec])}", file=sys.stderr, flush=True)
            f.close()
            continue
        u = f.get_u()[0]
        ip = f.get_ip_state()
        run = {"seconds": secs, "steps": st["n_steps"], "newton_iterations": st["n_iterations"],
               "pcg_iterations": st["n_pcg_iterations"], "ms_linear": st["ms_linear"], "ms_tangent": st["ms_tangent"],
               "plastic_fraction": float(np.mean(ip[:, f.ns + 2] == 0.0)), "dofs": f.ndof, "ips": f.nip}
        if ref is None:
            ref = (u, ip, f.ns)
        else:
            u0, ip0, ns = ref
            run["u_rel_err"] = float(np.abs(u - u0).max() / np.abs(u0).max())
            run["sigma_rel_err"] = float(np.abs(ip[:, :ns] - ip0[:, :ns]).max() / np.abs(ip0[:, :ns]).max())
            run["epbar_rel_err"] = float(np.abs(ip[:, ns + 3] - ip0[:, ns + 3]).max() / max(np.abs(ip0[:, ns + 3]).max(), 1e-300))
            run["plastic_flags_differ"] = int(np.sum((ip[:, ns + 2] == 0.0) != (ip0[:, ns + 2] == 0.0)))
        out["runs"][prec] = run
        print(f"{prec}: {json.dumps(run)}", file=sys.stderr, flush=True)
        f.close()
        del f, ip
    print(json.dumps(out), flush=True)


M_NAMES = {0: "Tri3", 1: "Tri6", 2: "Tri10", 3: "Tri15", 4: "Qua4", 5: "Qua8", 6: "Qua9", 7: "Tet4", 8: "Tet10", 9: "Hex8", 10: "Hex20"}


def run_pcg(a):
    """Time of one PCG iteration on the Hex20 tangent: one element pass assembles K, then linear solves
    with an unreachable tolerance run exactly --pcg-iters iterations each (the last one timed)."""
    from pmsim_amd.solver import Fem, PmError
    mesh, base, cfg, d = build_problem(a.nx, a.ny, a.nz, a.precision, a.kernel)
    cfg.set_pcg(1e-30, a.pcg_iters)
    f = Fem(d, cfg, keep_history=False)
    f.set_ddu(increment(mesh))
    f.backup_secondary_values()
    f.element_pass()
    b = np.random.default_rng(1).uniform(-1.0, 1.0, f.ndof)
    times = []
    for _ in range(max(1, a.warmup) + 1):
        t0 = time.time()
        try:
            f.linear_solve(b)
        except PmError:
            pass  # stops at the iteration limit by design
        times.append(time.time() - t0)
    ms = times[-1] / a.pcg_iters * 1e3
    print(json.dumps({"metric": "PCG ms per iteration (block Jacobi, K in the tangent's precision)", "value": ms,
                      "unit": "ms", "n_gpus": 1, "steps": a.pcg_iters, "warmup": max(1, a.warmup),
                      "higher_is_better": False, "dtype": a.precision, "data": "synthetic (Hex20 tangent after one pass)",
                      "config": {"workload": f"Hex20 {a.nx}x{a.ny}x{a.nz}, {f.ndof} DOFs, PCG on the free DOFs"}}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--nx", type=int, default=64)
    ap.add_argument("--ny", type=int, default=64)
    ap.add_argument("--nz", type=int, default=32)
    ap.add_argument("--precision", default="f64", choices=["f64", "f32", "bf16"])
    ap.add_argument("--kernel", default="default", choices=["default", "vector", "mfma", "tune"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workload", default="hex20", choices=["hex20", "mixed2d", "serve", "precision", "pcg"])
    ap.add_argument("--pcg-iters", type=int, default=1600, help="pcg: PCG iterations per timed solve")
    ap.add_argument("--n-inc", type=int, default=5, help="precision: increments of uz = -0.001")
    ap.add_argument("--precisions", default="f64,f32,bf16", help="precision: tangent precisions, FP64 first")
    ap.add_argument("--uz-stations", default="", help="precision: top uz after each increment (comma-separated) instead of --n-inc")
    ap.add_argument("--pcg-max", type=int, default=200000, help="precision: PCG iteration limit per linear solve")
    ap.add_argument("--requests", type=int, default=48, help="serve: number of requests")
    ap.add_argument("--rate", type=float, default=4.0, help="serve: Poisson arrival rate (requests/s)")
    ap.add_argument("--max-active", type=int, default=16, help="serve: concurrently running requests")
    ap.add_argument("--plain", action="store_true", help="serve: also time the plain FP64 form (one request at a time)")
    ap.add_argument("--hbm-gb", type=float, default=200.0, help="serve: device-memory budget for admission")
    ap.add_argument("--seed", type=int, default=20261203, help="serve: request and arrival seed")
    ap.add_argument("--elements", type=int, default=200_000, help="mixed2d: element count")
    ap.add_argument("--solve", action="store_true", help="mixed2d: also time the incremental Newton solve")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {a.gpus} is not supported (the element pass shards by mesh "
              f"partition; see the sharding design notes)", file=sys.stderr)
        sys.exit(2)
    if a.workload == "mixed2d":
        return run_mixed2d(a)
    if a.workload == "precision":
        return run_precision(a)
    if a.workload == "pcg":
        return run_pcg(a)
    if a.workload == "serve":
        # more hardware queues than the default 4 so the concurrent requests' streams do not serialise
        os.environ["GPU_MAX_HW_QUEUES"] = str(min(16, max(4, a.max_active)))
        return run_serve(a)
    from pmsim_amd.solver import Fem
    t0 = time.time()
    mesh, base, cfg, d = build_problem(a.nx, a.ny, a.nz, a.precision, a.kernel)
    f = Fem(d, cfg, keep_history=False)
    setup_s = time.time() - t0
    print(f"setup: {mesh.ncell} cells, {f.nip} IPs, {f.ndof} DOFs, {f.nnzb} blocks, {setup_s:.1f} s", file=sys.stderr, flush=True)
    # state == backup here, so this copy changes nothing; its known size (read + write of every IP
    # state array) calibrates the HBM byte counters of a profiled run
    f.backup_secondary_values()
    f.set_ddu(increment(mesh))
    for _ in range(a.warmup):
        f.element_pass()
    ms = f.time_element_pass(a.steps)
    st = f.get_ip_state()
    plastic = float(np.mean(st[:, f.ns + 2] == 0.0))
    nip = f.nip
    value = nip / (ms["pass_"] * 1e-3)
    tangent_flops = FLOP_PER_IP * nip
    achieved_tf = tangent_flops / (ms["tangent"] * 1e-3) / 1e12
    traffic, kname = measured_traffic(a.nx, a.ny, a.nz, a.precision, a.kernel)
    out = {
        "metric": "Hex20 IP-updates/s (K+R+return map)",
        "value": value,
        "unit": "IP-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms["pass_"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if a.precision == "f64" else f"{a.precision}-tangent/f64",
        "data": "synthetic (structured jittered Hex20 cube, seeded increment)",
        "config": {"workload": f"Hex20 unit cube {a.nx}x{a.ny}x{a.nz}, 27 IP, von Mises, one element pass",
                   "elements": int(mesh.ncell), "ips": int(nip), "dofs": int(f.ndof), "plastic_fraction": plastic,
                   "copy_calibration_bytes": int(2 * nip * (f.ns + 4) * 8)},
        "kernels_ms": {k: ms[k] for k in ("update", "residual", "tangent", "scatter", "total")},
        "roofline": {"bound": "mfma", "kernel": kname or "k_tangent", "achieved": achieved_tf, "peak": PEAK_TFLOPS[a.precision],
                     "unit": "TFLOP/s", "frac": achieved_tf / PEAK_TFLOPS[a.precision], "traffic": traffic,
                     "flop_per_ip": FLOP_PER_IP, "executed_flop_per_ip": EXECUTED_FLOP_PER_IP,
                     "executed_frac": EXECUTED_FLOP_PER_IP * nip / (ms["tangent"] * 1e-3) / 1e12 / PEAK_TFLOPS[a.precision]},
    }
    # the scatter is the HBM-bound half of the assembly: the stored element matrices in (upper node blocks,
    # 210 of the 400 per Hex20 element, 1,892 values with padding), the stored upper CSR blocks out (in the
    # tangent's precision), one map entry per element upper block and the map pointers
    nel = f.nip // 27
    ke_values = nel * (((20 * 21 // 2) * 9 + 3) // 4 * 4)
    vbytes = 8 if a.precision == "f64" else 4
    nnzu = (f.nnzb + f.ndof // 3) // 2
    scatter_bytes = ke_values * vbytes + nnzu * 9 * vbytes + nel * 210 * 4 + (nnzu + 1) * 4
    out["hbm"] = {"kernel": "k_scatter", "achieved": scatter_bytes / (ms["scatter"] * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                  "unit": "GB/s", "frac": scatter_bytes / (ms["scatter"] * 1e-3) / 1e9 / HBM_PEAK_GBS, "bytes": scatter_bytes}
    if not a.no_cpu:
        cb = cpu_baseline()
        out["cpu_baseline"] = cb
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
