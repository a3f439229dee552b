// Serving many independent solves on one GPU: footprint estimate, release of device state after a
// solve, and the admission scheduler (continuous batching over host worker threads, one HIP stream
// per request).  pmsim itself solves one problem per SolverImplicit::new + solve
// (src/fem/solver_implicit.rs:27-178); every request here runs exactly that path.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <numeric>
#include <thread>

#include "pm_internal.hpp"

namespace pm {

// Element passes of the running requests, launched together: each request's driver hands over the pass
// it needs (its stream idle) and waits; the engine takes every pass waiting at that moment, runs them as
// shared launches on its own stream (pass_batch_run) and releases their owners.
struct PassEngine {
  struct Item { PassJob job; bool done = false; int status = 0; };
  int device = 0;
  hipStream_t stream = nullptr;
  PassBatchBuf buf;
  std::mutex m;
  std::condition_variable cv_engine, cv_done;
  std::deque<Item*> pending;
  bool stop = false;
  std::thread th;
  int64_t n_batches = 0, n_jobs = 0;  // shared launches made, passes they carried

  int start(int dev) {
    device = dev;
    if (hipSetDevice(dev) != hipSuccess || hipStreamCreateWithFlags(&stream, hipStreamNonBlocking) != hipSuccess) return 1;
    th = std::thread([this] { loop(); });
    return 0;
  }
  void shutdown() {
    {
      std::lock_guard<std::mutex> g(m);
      stop = true;
    }
    cv_engine.notify_all();
    if (th.joinable()) th.join();
    if (buf.d) (void)hipFree(buf.d);
    if (stream) (void)hipStreamDestroy(stream);
  }
  int run(const PassJob& j) {
    Item it;
    it.job = j;
    std::unique_lock<std::mutex> lk(m);
    pending.push_back(&it);
    cv_engine.notify_all();
    cv_done.wait(lk, [&] { return it.done; });
    return it.status;
  }
  void loop() {
    (void)hipSetDevice(device);
    for (;;) {
      std::vector<Item*> batch;
      {
        std::unique_lock<std::mutex> lk(m);
        cv_engine.wait(lk, [&] { return stop || !pending.empty(); });
        if (pending.empty()) return;  // stop with nothing left
        batch.assign(pending.begin(), pending.end());
        pending.clear();
      }
      std::vector<PassJob> jobs;
      jobs.reserve(batch.size());
      for (Item* it : batch) jobs.push_back(it->job);
      const int rc = pass_batch_run(&buf, stream, jobs);
      {
        std::lock_guard<std::mutex> g(m);
        for (Item* it : batch) { it->status = rc; it->done = true; }
        n_batches++;
        n_jobs += (int64_t)batch.size();
      }
      cv_done.notify_all();
    }
  }
};

PassEngine* pass_engine_create(int device) {
  auto* e = new PassEngine();
  if (e->start(device)) { e->shutdown(); delete e; return nullptr; }
  return e;
}

void pass_engine_destroy(PassEngine* e) {
  if (!e) return;
  e->shutdown();
  delete e;
}

int pass_engine_run(PassEngine* e, const PassJob& job) { return e->run(job); }

}  // namespace pm

using namespace pm;

// A serving loop on one device: requests are queued at submission, admitted in order while fewer than
// max_active run and their footprint estimates fit the budget, and each runs pmsim's path (initialize,
// Newton solve, release) on a worker thread with its own stream; the element passes and the
// multi-workgroup PCG iterations of the running requests go through the run's two batching engines.
struct pm_server {
  using clk = std::chrono::steady_clock;
  struct Req { pm_fem* f = nullptr; int64_t bytes = 0; pm_request_stats st{}; int state = 0; /* 0 queued, 1 running, 2 done */ };
  int device = 0;
  int64_t budget = 0;
  int max_active = 1;
  PcgEngine* pcg = nullptr;
  PassEngine* pass = nullptr;
  std::mutex m;
  std::condition_variable cv;
  std::vector<std::unique_ptr<Req>> reqs;
  std::deque<int64_t> queue;
  int active = 0;
  int64_t active_bytes = 0;
  bool stop = false;
  std::thread sched;
  std::vector<std::thread> workers;
  clk::time_point t0 = clk::now();

  double now_s() const { return std::chrono::duration<double>(clk::now() - t0).count(); }

  void work(Req* rq) {
    Req& r = *rq;
    pm_fem* f = r.f;
    int32_t nconv = 0;
    int rc = pm_initialize(f, device);
    if (!rc) rc = pm_solve(f, &nconv);
    if (f->initialized && pm_release_device(f) && !rc) rc = 1;
    f->engine = nullptr;
    f->pass_engine = nullptr;
    std::lock_guard<std::mutex> g(m);
    r.st.status = rc;
    r.st.t_done = now_s();
    r.st.n_steps = f->stats.n_steps;
    r.st.n_iterations = f->stats.n_iterations;
    r.st.n_pcg_iterations = f->stats.n_pcg_iterations;
    r.state = 2;
    active--;
    active_bytes -= r.bytes;
    cv.notify_all();
  }

  void schedule() {
    std::unique_lock<std::mutex> lk(m);
    for (;;) {
      while (!queue.empty()) {
        Req& r = *reqs[queue.front()];
        if (active >= max_active || (active > 0 && active_bytes + r.bytes > budget)) break;
        queue.pop_front();
        active++;
        active_bytes += r.bytes;
        r.state = 1;
        r.st.t_admit = now_s();
        Req* rq = &r;
        workers.emplace_back([this, rq] { work(rq); });
      }
      if (stop && queue.empty()) return;
      cv.wait(lk);  // a submission or a completion
    }
  }
};

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
  b += nip * (b2 + 2) * 8;                        // IP geometry (J^-1, det J, radius)
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

int pm_server_create(pm_server** out, int32_t device, int64_t hbm_budget_bytes, int32_t max_active) {
  if (!out) return 1;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) { set_global_error("invalid HIP device"); return 1; }
  auto* s = new pm_server();
  s->device = device;
  s->budget = hbm_budget_bytes;
  s->max_active = std::max(1, (int)max_active);
  s->pcg = engine_create(device);
  s->pass = pass_engine_create(device);
  if (!s->pcg || !s->pass) {
    if (s->pcg) engine_destroy(s->pcg);
    if (s->pass) pass_engine_destroy(s->pass);
    delete s;
    set_global_error("cannot start the serving engines");
    return 1;
  }
  s->sched = std::thread([s] { s->schedule(); });
  *out = s;
  return 0;
}

int pm_server_submit(pm_server* s, pm_fem* f, int64_t* ticket) {
  if (!s || !f || !ticket) return 1;
  int64_t bytes = 0;
  if (f->initialized || f->released) { f->err = "a served request must be configured and not yet initialized"; return 1; }
  if (pm_estimate_device_bytes(f, &bytes)) { f->err = "cannot estimate the device footprint"; return 1; }
  f->engine = s->pcg;
  f->pass_engine = s->pass;
  std::lock_guard<std::mutex> g(s->m);
  if (s->stop) { f->engine = nullptr; f->pass_engine = nullptr; f->err = "the server is shutting down"; return 1; }
  auto r = std::make_unique<pm_server::Req>();
  r->f = f;
  r->bytes = bytes;
  r->st.t_arrival = s->now_s();
  r->st.status = -1;
  r->st.device_bytes = bytes;
  *ticket = (int64_t)s->reqs.size();
  s->reqs.push_back(std::move(r));
  s->queue.push_back(*ticket);
  s->cv.notify_all();
  return 0;
}

int pm_server_poll(pm_server* s, int64_t ticket, pm_request_stats* out) {
  if (!s) return -1;
  std::lock_guard<std::mutex> g(s->m);
  if (ticket < 0 || ticket >= (int64_t)s->reqs.size()) return -1;
  const auto& r = *s->reqs[ticket];
  if (r.state != 2) return 1;
  if (out) *out = r.st;
  return 0;
}

int pm_server_wait(pm_server* s, int64_t ticket, pm_request_stats* out) {
  if (!s) return -1;
  std::unique_lock<std::mutex> lk(s->m);
  if (ticket < 0 || ticket >= (int64_t)s->reqs.size()) return -1;
  s->cv.wait(lk, [&] { return s->reqs[ticket]->state == 2; });
  if (out) *out = s->reqs[ticket]->st;
  return 0;
}

int pm_server_stats(pm_server* s, int64_t* out) {
  if (!s || !out) return 1;
  std::lock_guard<std::mutex> g(s->pass->m);
  out[0] = s->pass->n_batches;
  out[1] = s->pass->n_jobs;
  return 0;
}

int pm_server_destroy(pm_server* s) {
  if (!s) return 0;
  {
    std::unique_lock<std::mutex> lk(s->m);
    s->stop = true;
    s->cv.notify_all();
  }
  if (s->sched.joinable()) s->sched.join();  // the queue is empty: every request was admitted
  for (auto& w : s->workers) w.join();
  engine_destroy(s->pcg);
  pass_engine_destroy(s->pass);
  delete s;
  return 0;
}

int pm_serve(int32_t n, pm_fem* const* fems, const double* arrival_s, int32_t device, int64_t hbm_budget_bytes,
             int32_t max_active, pm_request_stats* out) {
  if (n <= 0) return 0;
  using clk = std::chrono::steady_clock;
  for (int i = 0; i < n; i++) {
    int64_t b = 0;
    if (!fems[i] || pm_estimate_device_bytes(fems[i], &b)) return 1;
  }
  std::vector<int> order(n);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return arrival_s[a] < arrival_s[b]; });
  pm_server* s = nullptr;
  if (pm_server_create(&s, device, hbm_budget_bytes, max_active)) return 1;
  std::vector<int64_t> ticket(n, -1);
  int rc = 0;
  for (int i : order) {  // each request is submitted at its arrival time
    std::this_thread::sleep_until(s->t0 + std::chrono::duration_cast<clk::duration>(std::chrono::duration<double>(arrival_s[i])));
    if (pm_server_submit(s, fems[i], &ticket[i])) { rc = 1; break; }
  }
  for (int i = 0; i < n; i++) {
    out[i] = pm_request_stats{};
    out[i].status = -1;
    if (ticket[i] >= 0) pm_server_wait(s, ticket[i], &out[i]);
    out[i].t_arrival = arrival_s[i];
  }
  pm_server_destroy(s);
  return rc;
}

}  // extern "C"
