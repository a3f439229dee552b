"""Benchmark: Hex20 IP-updates/s of the implicit element pass (K + R + return map) on one MI355X.

Workload: unit cube, structured 64x64x32 Hex20 (131,072 elements, 27 IPs each = 3,538,944 IPs,
549,185 points, 1,647,555 DOFs), corner jitter ±5% of h (seed 20261203), von Mises E=2e5 nu=0.3
z0=250 H=1e3.  A step is one element pass over device-resident data: gather of coordinates and
displacement increments, strain increment, radial return from the backup state (with per-IP
sigma, z, lambda, loading flag, epbar), B^T sigma into the element forces and their deterministic
gather into R, B^T D B node blocks and their deterministic gather into the node-block CSR tangent.
The increment is a seeded field large enough that part of the IPs return plastically.

The same line carries `served`: the served half of the headline metric on a bounded seeded stream of
parametric strip-footing solves (Poisson arrivals near the service rate), as solves/s and p50/p99 latency
for the plain FP64 GPU form (one request at a time), the served FP64 form and the served mixed-precision
form (FP32 tangents), with the largest relative displacement difference of each against FP64.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--nx 64 --ny 64 --nz 32] [--no-cpu] [--no-serve]
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
FLOP_PER_IP = 2 * 60 * 60 * 6 + 2 * 6 * 6 * 60  # dense B^T D B per Hex20 IP = 47,520 (not what the kernel executes)


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


def _profile(nx, ny, nz, precision, kernel, what):
    """A committed rocprofv3 summary of the tangent kernel on this workload (JSON beside its .md), or None."""
    if (nx, ny, nz) != (64, 64, 32):
        return None
    tag = {("f64", "default"): "f64_vector", ("f64", "vector"): "f64_vector", ("f64", "mfma"): "f64_mfma",
           ("f32", "default"): "f32_vector", ("f32", "vector"): "f32_vector", ("f32", "mfma"): "f32_mfma",
           ("bf16", "default"): "bf16_mfma", ("bf16", "mfma"): "bf16_mfma"}.get((precision, kernel))
    path = os.path.join(ROOT, "profiles", f"{what}_hex20_64x64x32_{tag}.json")
    if tag is None or not os.path.exists(path):
        return None
    return json.load(open(path))


def measured_traffic(nx, ny, nz, precision, kernel):
    """HBM bytes per launch of the tangent kernel from the committed FETCH_SIZE / WRITE_SIZE summary of this
    workload (scaled by the copy-kernel calibration stored with it), or None."""
    prof = _profile(nx, ny, nz, precision, kernel, "element_pass")
    if prof is None:
        return None, None
    for name, k in prof["kernels"].items():
        if name.startswith("k_tangent"):
            return k["read_bytes"] + k["write_bytes"], name
    return None, None


def measured_ops(nx, ny, nz, precision, kernel):
    """Operations the tangent kernel executes per launch, from the committed SQ counter summary of this
    workload: FP64 VALU flops (SQ_INSTS_VALU_FLOPS_FP64: an FMA counts 2, per active lane) for the vector
    kernel, MFMA operations (SQ_INSTS_VALU_MFMA_MOPS_* x 512) for the matrix-core kernels.  None if absent."""
    prof = _profile(nx, ny, nz, precision, kernel, "k_tangent_counters")
    if prof is None:
        return None
    return prof


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


def cpu_baseline(seconds=12.0, seconds_all=6.0, n=(32, 32, 16)):
    """Oracle element pass (update + f_int + K per element) over a bounded sample of the same workload: on one
    host core (the reported value) and, as `all_cores`, with its OpenMP element loops on every core this process
    may use.  The sample (32x32x16 Hex20: 16,384 elements, 442,368 IPs, 35 MB of IP state and its backup) is
    larger than one CCD's L3; the rate does not depend on the sample size (compute-bound).  `full_solve` times
    the 16x8 Qua8 strip footing (20 increments) in the oracle beside the GPU solve of the same problem, and
    `hex20_solve` the Hex20 cube's loading on a 4x4x2 mesh likewise."""
    from oracle import oracle as orc
    mesh, base, cfg, d = build_problem(*n)
    o = orc.Oracle(d)
    o.init_state()
    ddu = increment(mesh)
    t1 = o.time_element_pass(ddu, reps=1, nthreads=1)
    reps = max(1, int(seconds / max(t1, 1e-6)))
    t = o.time_element_pass(ddu, reps=reps, nthreads=1) if reps > 1 else t1
    nip = mesh.ncell * 27
    tag = f"{n[0]}x{n[1]}x{n[2]} Hex20 ({mesh.ncell} elements, {nip} IPs)"
    out = dict(value=nip * reps / t, unit="IP-updates/s", cores=1, kind="port",
               sample=f"{tag} x {reps} passes, serial C++ oracle (update+f_int+K)", host=_host_cpu_model())
    nt = _host_threads()
    if nt > 1:
        t1 = o.time_element_pass(ddu, reps=1, nthreads=nt)
        reps2 = max(1, int(seconds_all / max(t1, 1e-6)))
        t2 = o.time_element_pass(ddu, reps=reps2, nthreads=nt)
        out["all_cores"] = dict(value=nip * reps2 / t2, unit="IP-updates/s", cores=nt,
                                sample=f"{tag} x {reps2} passes, C++ oracle with OpenMP element loops on {nt} threads")
    out["full_solve"] = full_solve_baseline()
    out["hex20_solve"] = hex20_solve_baseline()
    return out


def hex20_solve_baseline(n=(4, 4, 2)):
    """The Hex20 cube of the element-pass workload at a size the oracle solves in seconds (4x4x2, the same
    material and loading: top uz to -0.01 in 10 increments, 27 IPs, plastic almost everywhere at the end): the
    serial oracle (banded LU on the literal LMM system) against the GPU solve (FP64, PCG) of the same problem."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import cases
    from oracle import oracle as orc
    from pmsim_amd import describe
    from pmsim_amd.solver import Fem
    mesh, base, cfg, ess, nat = cases.hex20_cube(*n, n_inc=10, uz=-0.01)
    cfg.set_n_max_iterations(60)
    d = describe(mesh, base, cfg, ess, nat)
    t0 = time.time()
    uo, _, its = orc.Oracle(d).solve()
    cpu_s = time.time() - t0
    cfg.set_lin_sol("pcg")
    d = describe(mesh, base, cfg, ess, nat)
    Fem(d, cfg, keep_history=False).solve()  # warm-up (code objects, graphs)
    g = Fem(d, cfg, keep_history=False)
    t0 = time.time()
    g.solve()
    gpu_s = time.time() - t0
    u = g.get_u()[0]
    return {"workload": f"Hex20 cube {n[0]}x{n[1]}x{n[2]}, 27 IP, von Mises, top uz to -0.01 in 10 increments",
            "cpu_s": cpu_s, "cpu_cores": 1, "gpu_s": gpu_s, "newton_cpu": int(np.abs(its).sum()),
            "newton_gpu": g.stats()["n_iterations"],
            "u_rel_diff": float(np.abs(u - uo[-1][: g.ndof]).max() / np.abs(uo[-1][: g.ndof]).max())}


def full_solve_baseline():
    """The 16x8 Qua8 strip footing on von Mises material (20 increments, LMM, Newton with the consistent tangent):
    the serial oracle (banded LU on the literal LMM system) against the GPU solve (PCG) of the same problem."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import cases
    from oracle import oracle as orc
    from pmsim_amd import describe
    from pmsim_amd.solver import Fem
    mesh, base, cfg, ess, nat = cases.footing_c1()
    d = describe(mesh, base, cfg, ess, nat)
    t0 = time.time()
    uo, _, its = orc.Oracle(d).solve()
    cpu_s = time.time() - t0
    f = Fem(d, cfg, keep_history=False)
    f.solve()  # warm-up (code objects, graphs)
    g = Fem(d, cfg, keep_history=False)
    t0 = time.time()
    g.solve()
    gpu_s = time.time() - t0
    u = g.get_u()[0]
    return {"workload": "16x8 Qua8 strip footing, 3x3 IP, von Mises, 20 increments", "cpu_s": cpu_s, "cpu_cores": 1,
            "gpu_s": gpu_s, "newton_cpu": int(np.abs(its).sum()), "newton_gpu": g.stats()["n_iterations"],
            "u_rel_diff": float(np.abs(u - uo[-1][: g.ndof]).max() / np.abs(uo[-1][: g.ndof]).max())}


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


def serve_comparison(n_req, rate, seed, max_active=16, hbm_gb=200.0, plain=True, verbose=True):
    """Serving: Poisson arrivals of parametric strip-footing solves, admitted by HBM footprint and solved
    concurrently (pm_serve): solves/s and latency percentiles of the served FP64 and served FP32-tangent runs of
    the same trace, optionally the plain FP64 GPU form (the same requests one at a time in arrival order), and
    the largest relative differences of the final displacements against the served FP64 ones."""
    from pmsim_amd import describe
    from pmsim_amd.solver import Fem, serve
    from pmsim_amd.workloads import footing_request, poisson_arrivals
    rng = np.random.default_rng(seed)
    reqs = []
    for _ in range(n_req):
        (mesh, base, cfg, ess, nat), p = footing_request(rng)
        reqs.append((mesh, base, cfg, ess, nat, p))
    arrivals = poisson_arrivals(np.random.default_rng(seed + 1), n_req, rate)
    if verbose:
        print(f"serve: {n_req} requests, {sum(r[0].ncell for r in reqs)} elements in all, arrivals over "
              f"{arrivals[-1]:.1f} s", file=sys.stderr, flush=True)
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
        res = serve(hs, arrivals, hbm_budget_bytes=int(hbm_gb * (1 << 30)), max_active=max_active)
        ok = [r for r in res if r["status"] == 0]
        lat = np.array([r["t_done"] - r["t_arrival"] for r in ok])
        span = max(r["t_done"] for r in res) - min(arrivals)
        runs[prec] = dict(u=[h.get_u()[0] if r["status"] == 0 else None for h, r in zip(hs, res)],
                          solves_per_s=len(ok) / span, p50=float(np.percentile(lat, 50)), p99=float(np.percentile(lat, 99)),
                          failed=len(res) - len(ok), newton=sum(r["n_iterations"] for r in ok),
                          pcg=sum(r["n_pcg_iterations"] for r in ok))
        if verbose:
            print(f"serve {prec}: {len(ok)}/{len(res)} solved in {span:.1f} s, p50 {runs[prec]['p50']:.2f} s, "
                  f"p99 {runs[prec]['p99']:.2f} s", file=sys.stderr, flush=True)
        del hs
    def rel(a, b):
        return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))
    du = [rel(u2, u1) for u1, u2 in zip(runs["f64"]["u"], runs["f32"]["u"]) if u1 is not None and u2 is not None]
    keys = ("solves_per_s", "p50", "p99", "failed", "newton", "pcg")
    out = {"requests": n_req, "rate_per_s": rate, "seed": seed, "max_active": max_active,
           "elements": [int(r[0].ncell) for r in reqs],
           "workload": f"{n_req} strip-footing solves (Qua8 nx in [8,256], ny in [4,128], 5-40 increments), "
                       f"Poisson arrivals at {rate}/s",
           "served_fp64": {k: runs["f64"][k] for k in keys},
           "served_fp32_tangent": {k: runs["f32"][k] for k in keys},
           "u_rel_diff_fp32_tangent_vs_fp64": {"max": max(du) if du else None, "median": float(np.median(du)) if du else None}}
    if plain:
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
            sec = time.time() - t1
            start = max(t_free, float(arrivals[i]))
            t_free = start + sec
            lat.append(t_free - float(arrivals[i]))
            solve_s.append(sec)
            if runs["f64"]["u"][i] is not None:
                up.append(rel(u, runs["f64"]["u"][i]))
            if verbose:
                print(f"plain {len(solve_s)}/{n_req}: {mesh.ncell} cells, {sec:.2f} s", file=sys.stderr, flush=True)
        span = t_free - float(np.min(arrivals))
        out["plain_fp64"] = {"solves_per_s": n_req / span, "p50": float(np.percentile(lat, 50)),
                             "p99": float(np.percentile(lat, 99)), "sum_solve_s": float(np.sum(solve_s)),
                             "u_rel_diff_vs_served_fp64_max": max(up) if up else None}
        out["speedup_vs_plain_fp64"] = {"served_fp64": runs["f64"]["solves_per_s"] / out["plain_fp64"]["solves_per_s"],
                                        "served_fp32_tangent": runs["f32"]["solves_per_s"] / out["plain_fp64"]["solves_per_s"]}
    return out


def run_serve(a):
    """`--workload serve`: the serving comparison as its own JSON line."""
    sv = serve_comparison(a.requests, a.rate, a.seed, a.max_active, a.hbm_gb, plain=a.plain)
    f64 = sv["served_fp64"]
    out = {"metric": "served solves/s", "value": f64["solves_per_s"], "unit": "solves/s", "n_gpus": 1, "steps": a.requests,
           "warmup": 0, "ms_per_step": 1e3 / f64["solves_per_s"], "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded parametric strip footings, Poisson arrivals)",
           "config": {"workload": sv["workload"], "max_active": a.max_active, "hbm_budget_gb": a.hbm_gb},
           "latency_s": {"p50": f64["p50"], "p99": f64["p99"]}, "served": sv}
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
        # a suffix picks the tangent kernel: "-vector" (FP64 products rounded when stored; the FP32 default) or
        # "-mfma" (matrix cores; the BF16 default)
        kern = prec.split("-")[1] if "-" in prec else "default"
        if kern not in ("default", "vector", "mfma"):
            raise SystemExit(f"unknown tangent kernel suffix in {prec!r}")
        cfg.set_tangent_kernel(kern)
        cfg.set_tangent_precision(prec.split("-")[0]).set_scaled_rr(1e-9).set_pcg(1e-9, a.pcg_max).set_n_max_iterations(30).set_verbose(True, True)
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
    ap.add_argument("--kernel", default="default", choices=["default", "vector", "vector-wg", "mfma", "mfma-wg", "tune"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serve", action="store_true", help="hex20: leave out the served comparison")
    ap.add_argument("--serve-requests", type=int, default=24, help="hex20: requests of the served comparison")
    ap.add_argument("--serve-rate", type=float, default=3.0, help="hex20: Poisson rate of the served comparison (1/s)")
    ap.add_argument("--workload", default="hex20", choices=["hex20", "mixed2d", "serve", "precision", "pcg"])
    ap.add_argument("--pcg-iters", type=int, default=1600, help="pcg: PCG iterations per timed solve")
    ap.add_argument("--n-inc", type=int, default=5, help="precision: increments of uz = -0.001")
    ap.add_argument("--precisions", default="f64,f32,bf16", help="precision: tangent precisions, FP64 first (a suffix -vector or -mfma picks the kernel)")
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
    # more hardware queues than the default 4 so that the concurrent served requests' streams do not
    # serialise (set before the first HIP call of the process)
    if a.workload == "serve" or (a.workload == "hex20" and not a.no_serve):
        os.environ["GPU_MAX_HW_QUEUES"] = str(min(16, max(4, a.max_active)))
    if a.workload == "serve":
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
    traffic, kname = measured_traffic(a.nx, a.ny, a.nz, a.precision, a.kernel)
    ops = measured_ops(a.nx, a.ny, a.nz, a.precision, a.kernel)
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
    }
    # roofline of the dominant kernel (the tangent) against the unit it runs on, with the operations it executes
    # as counted by the committed SQ counter summary of this workload; the dense B^T D B count is kept apart
    dense_tf = FLOP_PER_IP * nip / (ms["tangent"] * 1e-3) / 1e12
    if ops is not None:
        unit_peak = PEAK_TFLOPS["f64"] if ops["bound"] == "valu-f64" else PEAK_TFLOPS[a.precision]
        achieved = ops["ops_per_launch"] / (ms["tangent"] * 1e-3) / 1e12
        out["roofline"] = {"bound": ops["bound"], "kernel": ops["kernel"], "achieved": achieved, "peak": unit_peak,
                           "unit": "TFLOP/s", "frac": achieved / unit_peak, "traffic": traffic,
                           "ops_per_launch": ops["ops_per_launch"], "ops_per_ip": ops["ops_per_launch"] / nip,
                           "ops_counter": ops["counter"],
                           "dense_count": {"flop_per_ip": FLOP_PER_IP, "tflops_equivalent": dense_tf}}
    else:
        out["roofline"] = {"bound": None, "kernel": kname or "k_tangent", "achieved": None, "peak": None, "unit": "TFLOP/s",
                           "frac": None, "traffic": traffic, "dense_count": {"flop_per_ip": FLOP_PER_IP, "tflops_equivalent": dense_tf}}
    # the scatter is the HBM-bound half of the assembly: the stored element matrices in (upper node blocks,
    # 210 of the 400 per Hex20 element, 1,892 values with padding), the stored upper CSR blocks out (in the
    # tangent's precision), one map entry per element upper block and the map pointers
    nel = f.nip // 27
    ke_values = nel * (((20 * 21 // 2) * 9 + 3) // 4 * 4)
    vbytes = 8 if a.precision == "f64" else 4
    nnzu = (f.nnzb + f.ndof // 3) // 2
    scatter_bytes = ke_values * vbytes + nnzu * 9 * vbytes + nel * 210 * 4 + (nnzu + 1) * 4
    out["hbm"] = {"kernel": "k_scatter_w", "achieved": scatter_bytes / (ms["scatter"] * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                  "unit": "GB/s", "frac": scatter_bytes / (ms["scatter"] * 1e-3) / 1e9 / HBM_PEAK_GBS, "bytes": scatter_bytes}
    # the update per IP: the state from the backup in and the new state out (sigma, z, lambda, epbar read and
    # written, the flag written: 152 B), the stored geometry in (88 B), its share of the element's
    # displacement increments in and of its force vector out (60 x 8 B / 27 each)
    update_bytes = f.nip * (152 + 88 + 2 * 60 * 8 / 27)
    out["hbm_update"] = {"kernel": "k_update", "achieved": update_bytes / (ms["update"] * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": update_bytes / (ms["update"] * 1e-3) / 1e9 / HBM_PEAK_GBS, "bytes": update_bytes}
    f.close()
    del f
    if not a.no_serve:
        out["served"] = serve_comparison(a.serve_requests, a.serve_rate, a.seed, a.max_active, a.hbm_gb, plain=True)
    if not a.no_cpu:
        cb = cpu_baseline()
        out["cpu_baseline"] = cb
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
