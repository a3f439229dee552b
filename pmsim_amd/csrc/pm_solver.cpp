// Newton driver over device-resident state (SolverImplicit::solve / iterate,
// src/fem/solver_implicit.rs:56-272), with the Lagrange-multiplier system reproduced by
// condensation of the prescribed DOFs so that the free block can be solved by PCG:
//
//   [K  A^T] [x]   [r]      x_p = g = u_p - c,   K_ff x_f = r_f - K_fp g,   y = r_p - (K x)_p,
//   [A   0 ] [y] = [g]      u -= x, lambda -= y  (bc_prescribed.rs:83-134, solver_data.rs:135-152).
//
// The convergence norms are taken over the same augmented vectors as pmsim's, so iteration
// counts follow the reference's (convergence_control.rs:52-127).
//
// With the arc-length method (arc_length_control.rs, solver_implicit.rs:117-150,200-251) every
// iteration also solves K du1 = F_ext (augmented: the multiplier rows' right-hand side is zero), and
// the correction becomes -du = du2 - dl du1 with dl = (dg/du . du2 - g) / (dg/dl + dg/du . du1) for the
// constraint g = du.du + psi dl^2 F.F - ds^2 over the augmented vectors [u; lambda].
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>

#include "pm_internal.hpp"

namespace pm {

constexpr int64_t kHostDirectMaxDof = 300000;  // host envelope Cholesky is the fallback below this size

static void launch_pass(pm_fem* f, bool update, bool residual, bool tangent) {
  for (auto& g : f->groups) launch_update(f, g, update);
  if (residual) launch_residual_gather(f);
  if (tangent) {
    for (auto& g : f->groups) launch_tangent(f, g);
    launch_scatter(f);
  }
}

void release_graphs(pm_fem* f) {
  for (auto& ge : f->pass_graph)
    if (ge) { (void)hipGraphExecDestroy(ge); ge = nullptr; }
  if (f->pcg_graph) { (void)hipGraphExecDestroy(f->pcg_graph); f->pcg_graph = nullptr; }
  f->pcg_graph_x = nullptr;
}

// One element pass (every batch's update, then the gathers).  With use_graphs the launches of each
// (update, residual, tangent) combination are captured once into a HIP graph and replayed: one submit
// instead of 2 + 2 x (element batches) kernel launches.  The captured kernels read only buffers whose
// addresses are fixed for the handle's lifetime; tuning or a configuration change drops the graphs.
int element_pass(pm_fem* f, bool update, bool residual, bool tangent) {
  if (gp_refuse(f, update, tangent)) return 1;
  if (!f->cfg.use_graphs) {
    launch_pass(f, update, residual, tangent);
    if (hipGetLastError() != hipSuccess) { f->err = "kernel launch failed"; return 1; }
    return 0;
  }
  hipGraphExec_t& ge = f->pass_graph[(update ? 1 : 0) | (residual ? 2 : 0) | (tangent ? 4 : 0)];
  if (!ge) {
    hipGraph_t graph = nullptr;
    PM_HIP(hipStreamBeginCapture(f->stream, hipStreamCaptureModeRelaxed));
    launch_pass(f, update, residual, tangent);
    const hipError_t launched = hipGetLastError();
    PM_HIP(hipStreamEndCapture(f->stream, &graph));
    if (launched != hipSuccess) { (void)hipGraphDestroy(graph); f->err = "kernel launch failed"; return 1; }
    const hipError_t e = hipGraphInstantiate(&ge, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) { ge = nullptr; f->err = std::string("graph instantiation failed: ") + hipGetErrorString(e); return 1; }
  }
  PM_HIP(hipGraphLaunch(ge, f->stream));
  return 0;
}

// The pieces of an element pass the driver asks for: the update kernel (stress update from the backup when
// do_update, else f_e of the current state), the residual gather, the tangent and its gather.  Served
// requests hand them to the run's pass engine, which launches the passes of all running requests together;
// otherwise they run on the handle's stream (the update combinations as captured graphs).
int run_pass(pm_fem* f, bool update, bool do_update, bool residual, bool tangent) {
  if (gp_refuse(f, update && do_update, tangent)) return 1;
  if (f->pass_engine) {
    PM_HIP(hipStreamSynchronize(f->stream));  // the engine's stream reads what this stream wrote
    PassJob job;
    job.f = f; job.update = update; job.do_update = do_update; job.residual = residual; job.tangent = tangent;
    if (pass_engine_run(f->pass_engine, job)) { f->err = "batched element pass failed"; return 1; }
    return 0;
  }
  if (update) return element_pass(f, do_update, residual, tangent);
  if (residual) launch_residual_gather(f);
  if (tangent) {
    for (auto& g : f->groups) launch_tangent(f, g);
    launch_scatter(f);
  }
  return 0;
}

int copy_ip_state_to_host(const pm_fem* f, double* v) {
  const int nv = f->nsig + 4;
  for (const auto& g : f->groups) {
    const int64_t nipg = g.nel * g.nip;
    std::vector<double> h((size_t)nipg * nv);
    if (hipMemcpyAsync(h.data(), g.d_state, h.size() * sizeof(double), hipMemcpyDeviceToHost, f->stream) != hipSuccess) return 1;
    if (hipStreamSynchronize(f->stream) != hipSuccess) return 1;
    for (int64_t el = 0; el < g.nel; el++) {
      const int64_t ip0 = f->cell_ip[g.cells[el]];
      for (int q = 0; q < g.nip; q++)
        for (int c = 0; c < nv; c++) v[(ip0 + q) * nv + c] = h[(size_t)c * nipg + el * g.nip + q];
    }
  }
  return 0;
}

int copy_ip_state_to_device(pm_fem* f, const double* v) {
  const int nv = f->nsig + 4;
  for (auto& g : f->groups) {
    const int64_t nipg = g.nel * g.nip;
    std::vector<double> h((size_t)nipg * nv);
    for (int64_t el = 0; el < g.nel; el++) {
      const int64_t ip0 = f->cell_ip[g.cells[el]];
      for (int q = 0; q < g.nip; q++)
        for (int c = 0; c < nv; c++) h[(size_t)c * nipg + el * g.nip + q] = v[(ip0 + q) * nv + c];
    }
    PM_HIP(hipMemcpyAsync(g.d_state, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, f->stream));
    PM_HIP(hipMemcpyAsync(g.d_backup, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, f->stream));
    PM_HIP(hipStreamSynchronize(f->stream));
  }
  return 0;
}

// x = K_ff^-1 b_f (prescribed entries of x are zero); PCG with host direct fallback
int linear_solve_free(pm_fem* f, const double* d_b, double* d_x, int* iters) {
  *iters = 0;
  auto t0 = std::chrono::steady_clock::now();
  struct Tm {
    pm_fem* f; std::chrono::steady_clock::time_point t0;
    ~Tm() { f->stats.ms_linear += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count(); }
  } tm{f, t0};
  if (f->cfg.lin_sol == 0) {
    int rc = pcg_solve(f, d_b, d_x, iters);
    if (rc == 0) return 0;
    if (rc == 1) return 1;  // NaN/Inf: no fallback
    // 2: PCG did not converge, 3: its full-pattern K could not be allocated; both fall back below the size
    // the host envelope solver handles, else report the PCG failure
    if (f->ndof > kHostDirectMaxDof) return 1;
    f->stats.n_host_fallbacks++;
  }
  std::vector<double> b(f->ndof), x;
  PM_HIP(hipMemcpyAsync(b.data(), d_b, sizeof(double) * f->ndof, hipMemcpyDeviceToHost, f->stream));
  PM_HIP(hipStreamSynchronize(f->stream));
  for (int64_t i = 0; i < f->ndof; i++) if (f->is_prescribed[i]) b[i] = 0.0;
  if (host_direct_solve(f, b, x)) return 1;
  PM_HIP(hipMemcpyAsync(d_x, x.data(), sizeof(double) * f->ndof, hipMemcpyHostToDevice, f->stream));
  PM_HIP(hipStreamSynchronize(f->stream));
  return 0;
}

static int save_output(pm_fem* f) {
  if (!f->cfg.keep_history) return 0;
  std::vector<double> u(f->ndof), lam(f->nprescribed), ip((size_t)f->nip * (f->nsig + 4));
  PM_HIP(hipMemcpyAsync(u.data(), f->d_u, sizeof(double) * f->ndof, hipMemcpyDeviceToHost, f->stream));
  if (f->nprescribed) PM_HIP(hipMemcpyAsync(lam.data(), f->d_lam, sizeof(double) * f->nprescribed, hipMemcpyDeviceToHost, f->stream));
  PM_HIP(hipStreamSynchronize(f->stream));
  if (copy_ip_state_to_host(f, ip.data())) { f->err = "HIP error while copying the state"; return 1; }
  f->hist_u.push_back(std::move(u));
  f->hist_lam.push_back(std::move(lam));
  f->hist_ip.push_back(std::move(ip));
  f->hist_t.push_back(f->t);
  f->hist_ell.push_back(f->ell);
  return 0;
}

// external forces at the step's time when a natural load varies with time (always: at any rate)
static int set_external_forces(pm_fem* f, bool always = false) {
  if (!f->loads_vary && !always) return 0;
  assemble_f_ext(f, (int64_t)std::llround((f->t - f->cfg.t_ini) / f->cfg.dt));
  if (hipMemcpyAsync(f->d_fext, f->f_ext.data(), sizeof(double) * f->ndof, hipMemcpyHostToDevice, f->stream) != hipSuccess ||
      hipStreamSynchronize(f->stream) != hipSuccess) {
    f->err = "uploading the external forces failed";
    return 1;
  }
  return 0;
}

static void set_prescribed_values(pm_fem* f) {
  if (!f->nprescribed) return;
  std::vector<double> c(f->nprescribed);
  int64_t k = (int64_t)std::llround((f->t - f->cfg.t_ini) / f->cfg.dt);
  if (k < 0) k = 0;
  if (k >= f->nstation) k = f->nstation - 1;
  for (int64_t p = 0; p < f->nprescribed; p++) c[p] = f->ess_table[(size_t)p * f->nstation + k];
  (void)hipMemcpyAsync(f->d_cval, c.data(), sizeof(double) * f->nprescribed, hipMemcpyHostToDevice, f->stream);
  (void)hipStreamSynchronize(f->stream);
}

int newton_solve(pm_fem* f, int32_t* n_converged) {
  auto t_start = std::chrono::steady_clock::now();
  const pm_config& cfg = f->cfg;
  *n_converged = 0;
  f->stats.n_steps = 0; f->stats.n_iterations = 0; f->stats.n_pcg_iterations = 0; f->stats.n_host_fallbacks = 0;
  f->stats.ms_linear = 0.0;
  f->hist_u.clear(); f->hist_lam.clear(); f->hist_ip.clear(); f->hist_t.clear(); f->hist_ell.clear();
  // config.validate (config.rs:255-265), reported as solver_data.rs:45-47 does
  if (cfg.arc_length_psi < 0.0 || cfg.arc_length_psi > 1.0 || std::fabs(cfg.first_trial_loading_factor) < 1e-12) {
    f->err = "cannot allocate simulation because config.validate() failed";
    return 1;
  }
  if (!cfg.lagrange_mult_method)
    for (double v : f->ess_table) if (v != 0.0) { f->err = "the Lagrange multiplier method is required for non-zero prescribed values"; return 1; }
  const int64_t n = f->ndof, npr = f->nprescribed;
  const bool lmm = cfg.lagrange_mult_method != 0;
  const bool arc = cfg.arc_length_method != 0;
  const double psi = cfg.arc_length_psi;
  // time_control.initialize
  f->t = cfg.t_ini;
  f->ell = 0.0;
  if (cfg.dt < 1e-10) { f->err = "Δt is smaller than the allowed minimum"; return 1; }
  if (set_external_forces(f, true)) return 1;  // F_ext (and F.F) under the current configuration
  if (arc && arc_alloc(f)) return 1;
  // initialize_internal_values and fresh primary variables
  for (auto& g : f->groups) launch_init_state(f, g);
  dev_fill(f, f->d_u, 0.0, n);
  dev_fill(f, f->d_ddu, 0.0, n);
  if (npr) dev_fill(f, f->d_lam, 0.0, npr);
  if (run_pass(f, true, false, false, false)) return 1;  // f_e of the initial state
  if (save_output(f)) return 1;
  double t_out = f->t + cfg.dt_out;
  // work vectors
  double* g = f->d_pcg + 4 * n;      // prescribed residual rows (npr)
  double* tv = g + std::max<int64_t>(npr, 1);
  double* kt = tv + n;
  double* b = kt + n;
  double* kx = b + n;                // reuses b region only after b is consumed
  // arc-length control: du1 = K^-1 F_ext, the previous two converged [u; lambda], the step's dlambda
  const int64_t P = std::max<int64_t>(npr, 1);
  double* u_old = f->d_arc;
  double* u_anc = arc ? u_old + n : nullptr;
  double* du1 = arc ? u_old + 2 * n : nullptr;
  double* fb = arc ? u_old + 3 * n : nullptr;
  double* lam_old = arc ? u_old + 4 * n : nullptr;
  double* lam_anc = arc ? lam_old + P : nullptr;
  double* dlam = arc ? lam_old + 2 * P : nullptr;
  double* y1 = arc ? lam_old + 3 * P : nullptr;
  double dds = 0, dds_old = 0, dds_min = 0, dds_max = 0, ell_anc = 0, ell_old = 0, ddl = 0, dg_dl = 0;
  bool converged_old = false;
  if (arc) { dev_fill(f, u_old, 0.0, 2 * n); dev_fill(f, lam_old, 0.0, 4 * P); }
  auto aug_dot = [&](const double* a, const double* b_, const double* al, const double* bl) {
    double v = dev_dot(f, a, b_, n);
    if (lmm && npr) v += dev_dot(f, al, bl, npr);
    return v;
  };
  int step_n_failure = 0;
  bool converged = false;
  for (int timestep = 0; timestep < cfg.n_max_time_steps; timestep++) {
    if (f->t + cfg.dt > cfg.t_fin) { if (save_output(f)) return 1; break; }
    f->t += cfg.dt;
    set_prescribed_values(f);
    if (set_external_forces(f)) return 1;
    if (arc) {
      // trial_increments (arc_length_control.rs:56-83)
      double alpha = 0.0;
      if (timestep == 0) f->ell = cfg.first_trial_loading_factor;
      if (timestep > 0) {
        if (std::fabs(dds_old) < 1e-12) { f->err = "Δs_old is too small"; return 1; }
        alpha = dds / dds_old;
        f->ell = (1.0 + alpha) * ell_old - alpha * ell_anc;
      }
      dev_arc_trial(f, n, timestep > 0, alpha, u_old, u_anc, f->d_u, f->d_ddu);
      if (lmm) dev_arc_trial(f, npr, timestep > 0, alpha, lam_old, lam_anc, f->d_lam, dlam);
      ddl = f->ell - ell_old;
      converged_old = converged;
    } else {
      dev_fill(f, f->d_ddu, 0.0, n);
      f->ell = 1.0;
    }
    if (!cfg.linear_problem) for (auto& gr : f->groups) launch_reset_lambda(f, gr);
    converged = false;
    for (int it = 0; it < cfg.n_max_iterations; it++) {
      f->stats.n_iterations++;
      if (run_pass(f, false, false, true, false)) return 1;
      dev_residual(f, g, f->ell);
      double norm_rr = dev_max_abs(f, f->d_rr, n, nullptr);
      if (lmm && npr) norm_rr = std::max(norm_rr, dev_max_abs(f, g, npr, nullptr));
      // arc-length constraint and derivatives (arc_length_control.rs:86-105); |g| joins the norm
      double arc_g = 0.0;
      if (arc) {
        if (timestep > 0) {
          arc_g = aug_dot(f->d_ddu, f->d_ddu, dlam, dlam) + psi * ddl * ddl * f->ftf - dds * dds;
          dg_dl = 2.0 * psi * ddl * f->ftf;
        } else {
          dg_dl = 1.0;
        }
        if (!std::isfinite(arc_g)) { f->err = "Found NaN or Inf"; return 1; }
        norm_rr = std::max(norm_rr, std::fabs(arc_g));
      }
      if (!std::isfinite(norm_rr)) { f->err = "Found NaN or Inf"; return 1; }
      bool conv_rr = norm_rr < cfg.tol_rr_abs;
      if (!conv_rr && cfg.tol_rr_scaled > 0.0) conv_rr = norm_rr <= cfg.tol_rr_scaled * dev_max_abs(f, f->d_fint, n, nullptr);
      if (cfg.verbose & 2) {
        std::fprintf(stderr, "  step %d t=%g it %d max|R|=%.3e\n", timestep, f->t, it, norm_rr);
        std::fflush(stderr);
      }
      if (conv_rr) { converged = true; break; }
      if (it == 0 || !cfg.constant_tangent) {
        if (run_pass(f, false, false, false, true)) return 1;
        if (check_device_error(f)) return 1;
        if (cfg.lin_sol == 0) build_block_jacobi(f);
      }
      int pits = 0;
      if (arc) {
        // linear system I: K du1 = F_ext (solved into the correction buffer, whose PCG graph is kept)
        dev_masked(f, f->d_fext, fb);
        if (linear_solve_free(f, fb, f->d_mdu, &pits)) return 1;
        launch_copy(f, du1, f->d_mdu, n);
        if (lmm) dev_complete_correction(f, nullptr, du1, kx, y1, f->d_fext);
      }
      dev_rhs(f, g, tv, kt, b);
      int pits2 = 0;
      if (linear_solve_free(f, b, f->d_mdu, &pits2)) return 1;
      pits += pits2;
      if (lmm) dev_complete_correction(f, g, f->d_mdu, kx, f->d_y);
      double arc_dl = 0.0;
      if (arc) {
        // dl = (dg/du . du2 - g) / (dg/dl + dg/du . du1), dg/du = 2 du (zero at the first step);
        // -du = du2 - dl du1 (arc_length_control.rs:108-126)
        const double c1 = timestep > 0 ? 2.0 * aug_dot(f->d_ddu, du1, dlam, y1) : 0.0;
        const double c2 = timestep > 0 ? 2.0 * aug_dot(f->d_ddu, f->d_mdu, dlam, f->d_y) : 0.0;
        arc_dl = (c2 - arc_g) / (dg_dl + c1);
        dev_axpy(f, n, -arc_dl, du1, f->d_mdu);
        if (lmm) dev_axpy(f, npr, -arc_dl, y1, f->d_y);
      }
      // analyze_mdu over the augmented correction [x; y]
      bool conv_mdu = false;
      if (it == 0) {
        double nm = dev_max_abs(f, f->d_mdu, n, nullptr);
        if (lmm && npr) nm = std::max(nm, dev_max_abs(f, f->d_y, npr, nullptr));
        if (!std::isfinite(nm)) { f->err = "Found NaN or Inf in mdu"; return 1; }
        launch_copy(f, f->d_mdu0, f->d_mdu, n);
        if (lmm && npr) launch_copy(f, f->d_mdu0 + n, f->d_y, npr);
      } else {
        double rel = dev_max_rel(f, f->d_mdu, f->d_mdu0, n);
        if (lmm && npr) rel = std::max(rel, dev_max_rel(f, f->d_y, f->d_mdu0 + n, npr));
        if (!std::isfinite(rel)) { f->err = "Found NaN or Inf in mdu"; return 1; }
        conv_mdu = rel < cfg.tol_mdu_rel;
      }
      if (cfg.verbose & 2) {
        std::fprintf(stderr, "  step %d t=%g it %d pcg %d%s\n", timestep, f->t, it, pits, conv_mdu ? " (mdu converged)" : "");
        std::fflush(stderr);
      }
      if (conv_mdu) { converged = true; break; }
      dev_update_primary(f, f->d_mdu, f->d_y);
      if (arc) {
        if (lmm) dev_axpy(f, npr, -1.0, f->d_y, dlam);  // the multipliers' part of du
        f->ell += arc_dl;                              // update_load_factor (arc_length_control.rs:129-133)
        ddl += arc_dl;
      }
      // backup at the first iteration; later iterations update from the backup (= restore)
      if (it == 0 || cfg.linear_problem) for (auto& gr : f->groups) launch_copy(f, gr.d_backup, gr.d_state, gr.nel * gr.nip * (f->nsig + 4));
      if (run_pass(f, true, true, false, false)) return 1;
      if (check_device_error(f)) return 1;
      if (cfg.linear_problem) { converged = true; break; }
      if (!arc && it == cfg.n_max_iterations - 1) { f->err = "Newton-Raphson did not converge"; return 1; }
    }
    if (!converged && !arc) { f->err = "Newton-Raphson did not converge"; return 1; }
    if (converged) { (*n_converged)++; f->stats.n_steps++; }
    if (arc) {
      // step_adaptation (arc_length_control.rs:136-168)
      if (converged) {
        if (timestep == 0) {
          dds = std::sqrt(aug_dot(f->d_ddu, f->d_ddu, dlam, dlam) + psi * ddl * ddl * f->ftf);
          dds_max = dds;
          dds_min = dds / 1024.0;
        }
        dds_old = dds;
        if (converged_old) dds = std::min(std::max(2.0 * dds, dds_min), dds_max);
        launch_copy(f, u_anc, u_old, n);
        launch_copy(f, u_old, f->d_u, n);
        if (lmm && npr) { launch_copy(f, lam_anc, lam_old, npr); launch_copy(f, lam_old, f->d_lam, npr); }
        ell_anc = ell_old;
        ell_old = f->ell;
      } else {
        dds = converged_old ? std::max(dds / 2.0, dds_min) : std::max(dds / 4.0, dds_min);
      }
    }
    if ((cfg.verbose & 1) && converged) {
      std::fprintf(stderr, "step %d t=%g converged: %d Newton iterations in all, %lld PCG iterations in all\n", timestep, f->t,
                   f->stats.n_iterations, (long long)f->stats.n_pcg_iterations);
      std::fflush(stderr);
    }
    bool last = timestep == cfg.n_max_time_steps - 1;
    if ((converged && f->t >= t_out) || last) { if (save_output(f)) return 1; t_out += cfg.dt_out; }
    if (++step_n_failure > cfg.allowed_step_n_failure) { f->err = "too many step failures"; return 1; }
    if (f->t >= cfg.t_fin) break;
  }
  if (hipStreamSynchronize(f->stream) != hipSuccess) { f->err = "HIP error at the end of the solve"; return 1; }
  f->stats.ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
  return 0;
}

}  // namespace pm
