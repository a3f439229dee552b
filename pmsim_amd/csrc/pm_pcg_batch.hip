// Batched PCG for serving: one engine per pm_serve call runs the block-Jacobi PCG iterations of every
// served request that is currently in its linear solve as the same three kernel launches, instead of
// three launches per request per iteration.  Each request keeps its own matrix, vectors and partial-sum
// buffers; inside the batched kernels every request sees exactly the work split (workgroups, lanes per
// row, partial-sum order) of the single-request multi-workgroup PCG, so its iterates are bitwise the
// same whether it is solved alone or in a batch, and the same as pm_solve's.
//
// Requests hand a solve to the engine from their own worker thread (after synchronizing their stream)
// and block until it is done; the engine thread admits new systems at chunk boundaries (every 16
// iterations, where convergence is checked) and retires converged ones.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>

#include "pm_internal.hpp"
#include "pm_spmv.hpp"

namespace pm {

namespace {

constexpr int BTPB = 256;        // = the single-request PCG's workgroup size
constexpr int CHUNK = 16;        // iterations between convergence checks (even: rz slots alternate)
constexpr int MAX_SYS = 512;     // systems in one batch

struct PcgSys {
  int64_t np;
  const int32_t* frp;   // full-pattern copy of K (widened by the request before it hands the solve over)
  const int32_t* fcol;
  const void* kfull;
  int lpr;              // lanes per block row of the SpMV
  const double* freev;
  const double* binv;
  double *x, *r, *z, *p, *q;
  double *pa, *pb, *pc;   // partials: p.q (nq), r.z (nv), r.r (nv)
  double* rzs;            // rz slots [2]
  double* rr;
  int nq, nv;             // workgroups of the SpMV and of the vector kernels
  int nrun;               // iterations of the current chunk this system takes (fewer than CHUNK only just
                          // before its iteration limit, which the single-request PCG stops at exactly)
};

__device__ __forceinline__ double bsum(double v, double* sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s = BTPB / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

__device__ __forceinline__ double bsum_partials(const double* __restrict__ part, int n, double* sh) {
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) acc += part[i];
  return bsum(acc, sh);
}

template <int BS>
__device__ __forceinline__ void bj_apply(int64_t a, const double* __restrict__ binv, const double* __restrict__ fr, const double* rv,
                                         double* zv) {
#pragma unroll
  for (int i = 0; i < BS; i++) {
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < BS; j++) s += binv[a * BS * BS + i * BS + j] * rv[j];
    zv[i] = s * fr[a * BS + i];
  }
}

// map entry: system slot << 20 | workgroup index within the system
__device__ __forceinline__ void unpack(int32_t e, int& s, int& w) { s = e >> 20; w = e & ((1 << 20) - 1); }

template <int BS, typename VT>
__global__ __launch_bounds__(BTPB) void kb_q(const PcgSys* __restrict__ sys, const int32_t* __restrict__ map, int k) {
  __shared__ double sh[BTPB];
  int si, w;
  unpack(map[blockIdx.x], si, w);
  const PcgSys& S = sys[si];
  if (k >= S.nrun) return;  // uniform over the workgroup
  const int lpr = S.lpr;
  const int lane = (int)(threadIdx.x & (lpr - 1));
  const int64_t ngroups = (int64_t)S.nq * BTPB / lpr;
  double acc = 0.0;
  for (int64_t a = ((int64_t)w * BTPB + threadIdx.x) / lpr; a < S.np; a += ngroups) {
    double y[BS];
    full_row_lanes<BS, VT>(S.frp, S.fcol, static_cast<const VT*>(S.kfull), a, lane, lpr, S.p, y);
    if (lane == 0) {
#pragma unroll
      for (int i = 0; i < BS; i++) {
        const int64_t t = a * BS + i;
        const double sv = y[i] * S.freev[t];
        S.q[t] = sv;
        acc += S.p[t] * sv;
      }
    }
  }
  const double r = bsum(acc, sh);
  if (threadIdx.x == 0) S.pa[w] = r;
}

template <int BS>
__global__ __launch_bounds__(BTPB) void kb_xrz(const PcgSys* __restrict__ sys, const int32_t* __restrict__ map, int k) {
  __shared__ double sh[BTPB];
  int si, w;
  unpack(map[blockIdx.x], si, w);
  const PcgSys& S = sys[si];
  if (k >= S.nrun) return;  // uniform over the workgroup
  const int parity = k & 1;
  const double pq = bsum_partials(S.pa, S.nq, sh);
  const double alpha = pq != 0.0 ? S.rzs[parity] / pq : 0.0;
  double rz = 0.0, rr = 0.0;
  for (int64_t a = (int64_t)w * BTPB + threadIdx.x; a < S.np; a += (int64_t)S.nv * BTPB) {
    double rv[BS], zv[BS];
#pragma unroll
    for (int i = 0; i < BS; i++) {
      const int64_t k = a * BS + i;
      S.x[k] += alpha * S.p[k];
      rv[i] = S.r[k] - alpha * S.q[k];
      S.r[k] = rv[i];
      rr += rv[i] * rv[i];
    }
    bj_apply<BS>(a, S.binv, S.freev, rv, zv);
#pragma unroll
    for (int i = 0; i < BS; i++) { S.z[a * BS + i] = zv[i]; rz += rv[i] * zv[i]; }
  }
  const double s1 = bsum(rz, sh), s2 = bsum(rr, sh);
  if (threadIdx.x == 0) { S.pb[w] = s1; S.pc[w] = s2; }
}

template <int BS>
__global__ __launch_bounds__(BTPB) void kb_p(const PcgSys* __restrict__ sys, const int32_t* __restrict__ map, int k) {
  __shared__ double sh[BTPB];
  int si, w;
  unpack(map[blockIdx.x], si, w);
  const PcgSys& S = sys[si];
  if (k >= S.nrun) return;  // uniform over the workgroup
  const int parity = k & 1;
  const double rzn = bsum_partials(S.pb, S.nv, sh);
  const double rzo = S.rzs[parity];
  const double beta = rzo != 0.0 ? rzn / rzo : 0.0;
  if (w == 0) {
    const double rr = bsum_partials(S.pc, S.nv, sh);
    if (threadIdx.x == 0) { S.rzs[parity ^ 1] = rzn; *S.rr = rr; }
  }
  const int64_t n = S.np * BS;
  for (int64_t i = (int64_t)w * BTPB + threadIdx.x; i < n; i += (int64_t)S.nv * BTPB) S.p[i] = S.z[i] + beta * S.p[i];
}

// start of a solve (same arithmetic as the single-request start): r = b*free, x = 0, z = M^-1 r, p = z;
// partials of r.z and b.b over the system's vector workgroups
template <int BS>
__global__ __launch_bounds__(BTPB) void k_pcg_init_b(int64_t np, const double* __restrict__ b, const double* __restrict__ fr,
                                                     const double* __restrict__ binv, double* __restrict__ x, double* __restrict__ r,
                                                     double* __restrict__ z, double* __restrict__ p, double* __restrict__ part_rz,
                                                     double* __restrict__ part_bb) {
  __shared__ double sh[BTPB];
  double rz = 0.0, bb = 0.0;
  for (int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; a < np; a += (int64_t)gridDim.x * blockDim.x) {
    double rv[BS], zv[BS];
#pragma unroll
    for (int i = 0; i < BS; i++) { rv[i] = b[a * BS + i] * fr[a * BS + i]; r[a * BS + i] = rv[i]; x[a * BS + i] = 0.0; bb += rv[i] * rv[i]; }
    bj_apply<BS>(a, binv, fr, rv, zv);
#pragma unroll
    for (int i = 0; i < BS; i++) { z[a * BS + i] = zv[i]; p[a * BS + i] = zv[i]; rz += rv[i] * zv[i]; }
  }
  const double s1 = bsum(rz, sh), s2 = bsum(bb, sh);
  if (threadIdx.x == 0) { part_rz[blockIdx.x] = s1; part_bb[blockIdx.x] = s2; }
}

__global__ void k_pcg_init_fin_b(int npart, const double* __restrict__ part_rz, const double* __restrict__ part_bb, double* rz_out,
                                 double* bb_out) {
  __shared__ double sh[BTPB];
  const double a = bsum_partials(part_rz, npart, sh);
  const double b = bsum_partials(part_bb, npart, sh);
  if (threadIdx.x == 0) { *rz_out = a; *bb_out = b; }
}

__global__ void kb_gather(int n, const PcgSys* __restrict__ sys, double* __restrict__ out) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) out[i] = *sys[i].rr;
}

}  // namespace

struct PcgEngine {
  struct Job {
    pm_fem* f;
    const double* b;
    double* x;
    int bs = 2;
    bool lowp = false;
    double tol2 = 0.0;
    int max_it = 0, it = 0, status = 0;
    bool done = false;
    PcgSys desc;
  };
  int device = 0;
  hipStream_t stream = nullptr;
  std::mutex m;
  std::condition_variable cv_engine, cv_done;
  std::deque<Job*> pending;
  std::vector<Job*> active;
  bool stop = false;
  std::thread th;
  PcgSys* d_sys = nullptr;
  int32_t* d_map = nullptr;  // [class 0..3][q map | v map], each MAX_SYS*4096
  double* d_rr = nullptr;
  std::string err;

  static constexpr int64_t MAP_CAP = (int64_t)MAX_SYS * 4096;

  int start(int dev) {
    device = dev;
    if (hipSetDevice(dev) != hipSuccess || hipStreamCreateWithFlags(&stream, hipStreamNonBlocking) != hipSuccess) return 1;
    if (hipMalloc((void**)&d_sys, sizeof(PcgSys) * MAX_SYS) != hipSuccess) return 1;
    if (hipMalloc((void**)&d_map, sizeof(int32_t) * MAP_CAP * 2) != hipSuccess) return 1;
    if (hipMalloc((void**)&d_rr, sizeof(double) * MAX_SYS) != hipSuccess) return 1;
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
    if (d_sys) (void)hipFree(d_sys);
    if (d_map) (void)hipFree(d_map);
    if (d_rr) (void)hipFree(d_rr);
    if (stream) (void)hipStreamDestroy(stream);
  }

  // called from a request's worker thread; blocks until the system is solved
  int solve(pm_fem* f, const double* d_b, double* d_x, int* iters) {
    const int64_t np = f->npoint, n = np * f->bs;
    Job job;
    job.f = f; job.b = d_b; job.x = d_x; job.bs = f->bs;
    job.lowp = f->cfg.tangent_precision != 0 && f->d_kval32;
    job.max_it = std::max(0, f->cfg.pcg_max_iter);  // exactly the single-request limit
    PcgSys& S = job.desc;
    S.np = np; S.frp = f->d_frp; S.fcol = f->d_fcol;
    S.kfull = job.lowp ? (const void*)f->d_kfull32 : (const void*)f->d_kfull;
    S.lpr = lanes_per_row(f->nnzb, np);
    S.freev = f->d_free; S.binv = f->d_binv;
    S.r = f->d_pcg; S.z = S.r + n; S.p = S.z + n; S.q = S.p + n; S.x = d_x;
    S.nv = (int)std::min<int64_t>(2048, std::max<int64_t>(1, (np + BTPB - 1) / BTPB));
    S.nq = spmv_workgroups(np, S.lpr, BTPB);
    S.pa = f->d_red; S.pb = S.pa + S.nq; S.pc = S.pb + S.nv;
    S.rzs = f->d_scal + 16; S.rr = f->d_scal + 18;
    S.nrun = CHUNK;
    if (hipStreamSynchronize(f->stream) != hipSuccess) { f->err = "stream synchronization failed"; return 1; }
    std::unique_lock<std::mutex> lk(m);
    pending.push_back(&job);
    cv_engine.notify_all();
    cv_done.wait(lk, [&] { return job.done; });
    *iters = job.it;
    f->stats.n_pcg_iterations += job.it;
    if (job.status == 1) f->err = "Found NaN or Inf in mdu";
    if (job.status == 2) f->err = "PCG did not converge";
    if (job.status == 3) f->err = "batched PCG launch failed: " + err;
    return job.status;
  }

  void finish(Job* j, int status) {  // caller holds m
    j->status = status;
    j->done = true;
  }

  // start a system: r = b*free, x = 0, z = M^-1 r, p = z; rz and b.b.  Returns -1 when the system joins
  // the batch, else its final status (0: b = 0, 1: not finite, 3: device failure).
  int admit(Job* j) {
    pm_fem* f = j->f;
    const int64_t np = f->npoint;
    const PcgSys& S = j->desc;
    double* bb_s = f->d_scal + 19;
    if (j->bs == 2)
      hipLaunchKernelGGL((k_pcg_init_b<2>), dim3(S.nv), dim3(BTPB), 0, stream, np, j->b, S.freev, S.binv, S.x, S.r, S.z, S.p, S.pb, S.pc);
    else
      hipLaunchKernelGGL((k_pcg_init_b<3>), dim3(S.nv), dim3(BTPB), 0, stream, np, j->b, S.freev, S.binv, S.x, S.r, S.z, S.p, S.pb, S.pc);
    hipLaunchKernelGGL(k_pcg_init_fin_b, dim3(1), dim3(BTPB), 0, stream, S.nv, S.pb, S.pc, S.rzs, bb_s);
    double bb = 0.0;
    if (hipMemcpyAsync(&bb, bb_s, sizeof(double), hipMemcpyDeviceToHost, stream) != hipSuccess || hipStreamSynchronize(stream) != hipSuccess) {
      err = "initial residual";
      return 3;
    }
    if (bb == 0.0) return 0;
    if (!std::isfinite(bb)) return 1;
    j->tol2 = f->cfg.pcg_rtol * f->cfg.pcg_rtol * bb;
    if (j->max_it == 0) return bb <= j->tol2 ? 0 : 2;  // no iteration allowed: the start residual decides
    return -1;
  }

  void loop() {
    (void)hipSetDevice(device);
    std::vector<PcgSys> h_sys;
    std::vector<int32_t> h_map[4][2];  // class = (bs==3)*2 + lowp
    std::vector<double> h_rr;
    bool dirty = true;
    for (;;) {
      std::vector<Job*> fresh;
      {
        std::unique_lock<std::mutex> lk(m);
        cv_engine.wait(lk, [&] { return stop || !pending.empty() || !active.empty(); });
        if (stop && pending.empty() && active.empty()) return;
        while (!pending.empty() && (int)(active.size() + fresh.size()) < MAX_SYS) { fresh.push_back(pending.front()); pending.pop_front(); }
      }
      bool any_done = false;
      for (Job* j : fresh) {
        const int st = admit(j);
        if (st < 0) { active.push_back(j); dirty = true; continue; }
        std::lock_guard<std::mutex> g(m);
        finish(j, st);
        any_done = true;
      }
      if (any_done) cv_done.notify_all();
      if (active.empty()) continue;
      // a system close to its iteration limit takes only the iterations left (the single-request PCG
      // checks at every 16th iteration and at the limit itself)
      for (size_t s = 0; s < active.size(); s++) {
        const int left = std::min(CHUNK, active[s]->max_it - active[s]->it);
        if (active[s]->desc.nrun != left) { active[s]->desc.nrun = left; dirty = true; }
      }
      if (dirty) {
        h_sys.resize(active.size());
        for (auto& c : h_map) { c[0].clear(); c[1].clear(); }
        for (size_t s = 0; s < active.size(); s++) {
          h_sys[s] = active[s]->desc;
          const int cls = (active[s]->bs == 3 ? 2 : 0) + (active[s]->lowp ? 1 : 0);
          for (int w = 0; w < h_sys[s].nq; w++) h_map[cls][0].push_back((int32_t)(s << 20) | w);
          for (int w = 0; w < h_sys[s].nv; w++) h_map[cls][1].push_back((int32_t)(s << 20) | w);
        }
        (void)hipMemcpyAsync(d_sys, h_sys.data(), sizeof(PcgSys) * h_sys.size(), hipMemcpyHostToDevice, stream);
        int64_t off = 0;
        for (auto& c : h_map)
          for (auto& v : c) {
            if (!v.empty()) (void)hipMemcpyAsync(d_map + off, v.data(), sizeof(int32_t) * v.size(), hipMemcpyHostToDevice, stream);
            off += (int64_t)v.size();
          }
        (void)hipStreamSynchronize(stream);  // host tables may change before the next upload
        dirty = false;
      }
      for (int k = 0; k < CHUNK; k++) {
        int64_t off = 0;
        for (int cls = 0; cls < 4; cls++) {
          const int32_t* mq = d_map + off;
          off += (int64_t)h_map[cls][0].size();
          const int32_t* mv = d_map + off;
          off += (int64_t)h_map[cls][1].size();
          const unsigned gq = (unsigned)h_map[cls][0].size(), gv = (unsigned)h_map[cls][1].size();
          if (!gq) continue;
          if (cls == 0) hipLaunchKernelGGL((kb_q<2, double>), dim3(gq), dim3(BTPB), 0, stream, d_sys, mq, k);
          if (cls == 1) hipLaunchKernelGGL((kb_q<2, float>), dim3(gq), dim3(BTPB), 0, stream, d_sys, mq, k);
          if (cls == 2) hipLaunchKernelGGL((kb_q<3, double>), dim3(gq), dim3(BTPB), 0, stream, d_sys, mq, k);
          if (cls == 3) hipLaunchKernelGGL((kb_q<3, float>), dim3(gq), dim3(BTPB), 0, stream, d_sys, mq, k);
          if (cls < 2) {
            hipLaunchKernelGGL((kb_xrz<2>), dim3(gv), dim3(BTPB), 0, stream, d_sys, mv, k);
            hipLaunchKernelGGL((kb_p<2>), dim3(gv), dim3(BTPB), 0, stream, d_sys, mv, k);
          } else {
            hipLaunchKernelGGL((kb_xrz<3>), dim3(gv), dim3(BTPB), 0, stream, d_sys, mv, k);
            hipLaunchKernelGGL((kb_p<3>), dim3(gv), dim3(BTPB), 0, stream, d_sys, mv, k);
          }
        }
      }
      const int ns = (int)active.size();
      hipLaunchKernelGGL(kb_gather, dim3((ns + 255) / 256), dim3(256), 0, stream, ns, d_sys, d_rr);
      h_rr.resize(ns);
      const bool ok = hipMemcpyAsync(h_rr.data(), d_rr, sizeof(double) * ns, hipMemcpyDeviceToHost, stream) == hipSuccess &&
                      hipStreamSynchronize(stream) == hipSuccess && hipGetLastError() == hipSuccess;
      std::vector<Job*> keep;
      {
        std::lock_guard<std::mutex> g(m);
        for (int s = 0; s < ns; s++) {
          Job* j = active[s];
          j->it += j->desc.nrun;
          if (!ok) { err = "kernel execution"; finish(j, 3); }
          else if (!std::isfinite(h_rr[s])) finish(j, 1);
          else if (h_rr[s] <= j->tol2) finish(j, 0);
          else if (j->it >= j->max_it) finish(j, 2);
          else keep.push_back(j);
        }
      }
      if (keep.size() != active.size()) { active.swap(keep); dirty = true; cv_done.notify_all(); }
    }
  }
};

PcgEngine* engine_create(int device) {
  auto* e = new PcgEngine();
  if (e->start(device)) { e->shutdown(); delete e; return nullptr; }
  return e;
}

void engine_destroy(PcgEngine* e) {
  if (!e) return;
  e->shutdown();
  delete e;
}

int engine_solve(PcgEngine* e, pm_fem* f, const double* d_b, double* d_x, int* iters) { return e->solve(f, d_b, d_x, iters); }

}  // namespace pm
