// Serving many independent solves on one GPU: footprint estimate, release of device state after a
// solve, and the admission scheduler (continuous batching over host worker threads, one HIP stream
// per request).  pmsim itself solves one problem per SolverImplicit::new + solve
// (src/fem/solver_implicit.rs:27-178); every request here runs exactly that path.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <mutex>
#include <numeric>
#include <thread>

#include "pm_internal.hpp"

using namespace pm;

extern "C" {

int pm_estimate_device_bytes(const pm_fem* f, int64_t* bytes) {
  if (!f) return 1;
  const int nd = f->ndim, ns = f->nsig;
  const int64_t b2 = (int64_t)nd * nd;
  int64_t blocks = 0, kev = 0, nip = 0, nfe = 0, conn = 0;
  for (int64_t e = 0; e < f->ncell; e++) {
    const int k = f->kind[e];
    const int nn = kind_nnode(k);
    if (nn <= 0) return 1;
    auto it = f->params.find(f->attr[e]);
    const int ng = (it != f->params.end() && it->second.ngauss > 0) ? it->second.ngauss : kind_default_ngauss(k);
    blocks += (int64_t)nn * nn;
    kev += packed_ke_values(nn, nd);
    nip += ng;
    nfe += (int64_t)nn * nd;
    conn += nn;
  }
  const int64_t ndof = f->npoint * nd, np = std::max<int64_t>(1, (int64_t)f->ess_point.size());
  const int64_t ts = f->cfg.tangent_precision == 0 ? 8 : 4;
  int64_t b = 0;
  b += f->npoint * nd * 8;                       // coordinates
  b += conn * 5;                                  // connectivity + ranks
  b += nip * (ns + 4) * 8 * 2;                    // IP state + backup
  b += kev * ts;                                  // element matrices (packed upper blocks)
  b += blocks * (b2 * ts + 16);                   // K (upper blocks) in the tangent's precision, upper and lower index
                                                  // lists and the gather map (bound: every count <= element blocks)
  b += nfe * (8 + 4);                             // element forces + R map
  b += (f->npoint + 1) * 4 * 3;                   // row pointers
  b += ndof * 8 * 16 + np * 8 * 8;                // vectors, PCG work, multipliers
  if (f->cfg.arc_length_method) b += (4 * ndof + 4 * np) * 8;  // arc-length vectors
  if (ndof > f->pcg_small_max) b += blocks * (b2 * ts + 8) + (f->npoint + 1) * 4;  // full-pattern K copy (multi-workgroup PCG)
  b += f->npoint * b2 * 8;                        // block-Jacobi inverses
  b += 1 << 20;                                   // reductions, scalars, shape tables
  *bytes = b;
  return 0;
}

int pm_release_device(pm_fem* f) {
  if (!f) return 1;
  if (!f->initialized) { f->err = "pm_initialize has not been called"; return 1; }
  (void)hipSetDevice(f->device);
  f->rel_u.assign(f->ndof, 0.0);
  f->rel_lam.assign(f->nprescribed, 0.0);
  f->rel_ip.assign((size_t)f->nip * (f->nsig + 4), 0.0);
  int rc = 0;
  if (hipMemcpyAsync(f->rel_u.data(), f->d_u, sizeof(double) * f->ndof, hipMemcpyDeviceToHost, f->stream) != hipSuccess) rc = 1;
  if (f->nprescribed && hipMemcpyAsync(f->rel_lam.data(), f->d_lam, sizeof(double) * f->nprescribed, hipMemcpyDeviceToHost,
                                       f->stream) != hipSuccess)
    rc = 1;
  if (hipStreamSynchronize(f->stream) != hipSuccess) rc = 1;
  if (!rc) rc = copy_ip_state_to_host(f, f->rel_ip.data());
  device_free(f);
  f->initialized = false;
  f->released = true;
  if (rc) f->err = "copying results to the host failed";
  return rc;
}

int pm_serve(int32_t n, pm_fem* const* fems, const double* arrival_s, int32_t device, int64_t hbm_budget_bytes,
             int32_t max_active, pm_request_stats* out) {
  if (n <= 0) return 0;
  if (max_active < 1) max_active = 1;
  using clk = std::chrono::steady_clock;
  std::vector<int> order(n);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return arrival_s[a] < arrival_s[b]; });
  std::vector<int64_t> bytes(n, 0);
  for (int i = 0; i < n; i++) {
    out[i] = pm_request_stats{};
    out[i].t_arrival = arrival_s[i];
    out[i].status = -1;
    if (!fems[i] || pm_estimate_device_bytes(fems[i], &bytes[i])) return 1;
    out[i].device_bytes = bytes[i];
  }
  // one batched PCG engine for the run: every served request's multi-workgroup PCG iterations go
  // through it (systems small enough for the single-workgroup PCG keep that one-launch path)
  PcgEngine* engine = engine_create(device);
  if (!engine) return 1;
  for (int i = 0; i < n; i++) fems[i]->engine = engine;
  std::mutex m;
  std::condition_variable cv;
  int active = 0, done = 0;
  int64_t active_bytes = 0;
  std::vector<std::thread> workers;
  workers.reserve(n);
  const auto t0 = clk::now();
  auto now_s = [t0] { return std::chrono::duration<double>(clk::now() - t0).count(); };
  size_t next = 0;
  std::unique_lock<std::mutex> lk(m);
  while (done < n) {
    // admit in arrival order while the head of the queue has arrived and fits
    while (next < (size_t)n) {
      const int i = order[next];
      if (arrival_s[i] > now_s() || active >= max_active) break;
      if (active > 0 && active_bytes + bytes[i] > hbm_budget_bytes) break;
      active++;
      active_bytes += bytes[i];
      next++;
      out[i].t_admit = now_s();
      workers.emplace_back([&, i] {
        pm_fem* f = fems[i];
        int32_t nconv = 0;
        int rc = pm_initialize(f, device);
        if (!rc) rc = pm_solve(f, &nconv);
        if (f->initialized && pm_release_device(f) && !rc) rc = 1;
        std::lock_guard<std::mutex> g(m);
        out[i].status = rc;
        out[i].t_done = now_s();
        out[i].n_steps = f->stats.n_steps;
        out[i].n_iterations = f->stats.n_iterations;
        out[i].n_pcg_iterations = f->stats.n_pcg_iterations;
        active--;
        active_bytes -= bytes[i];
        done++;
        cv.notify_all();
      });
    }
    if (done == n) break;
    if (next < (size_t)n && active < max_active && (active == 0 || active_bytes + bytes[order[next]] <= hbm_budget_bytes)) {
      // only the arrival time holds the head back
      cv.wait_until(lk, t0 + std::chrono::duration_cast<clk::duration>(std::chrono::duration<double>(arrival_s[order[next]])));
    } else {
      cv.wait(lk);  // a completion frees capacity
    }
  }
  lk.unlock();
  for (auto& w : workers) w.join();
  engine_destroy(engine);
  for (int i = 0; i < n; i++) fems[i]->engine = nullptr;
  return 0;
}

}  // extern "C"
