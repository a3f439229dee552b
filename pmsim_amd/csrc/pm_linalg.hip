// Device linear algebra of the Newton driver: node-block CSR SpMV, block-Jacobi PCG (FP64 vectors,
// FP64 or FP32 matrix values), deterministic fixed-shape reductions, and the vector updates of
// pmsim's residual / multiplier / primary-variable steps (solver_data.rs:117-152, bc_prescribed.rs:99-108).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "pm_internal.hpp"

namespace pm {

namespace {

constexpr int TPB = 256;
constexpr int RED_BLOCKS = 1024;  // fixed partial count -> deterministic reduction order

enum Slot { S_RZ = 0, S_PQ, S_ALPHA, S_BETA, S_RR, S_BB, S_RZNEW, S_MAX, S_TMP, NSLOT };

__device__ __forceinline__ double block_sum(double v, double* sh) {
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (t < s) sh[t] += sh[t + s];
    __syncthreads();
  }
  double r = sh[0];
  __syncthreads();
  return r;
}

__device__ __forceinline__ double block_max(double v, double* sh) {
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (t < s) { double a = sh[t], b = sh[t + s]; sh[t] = (b > a || b != b) ? b : a; }
    __syncthreads();
  }
  double r = sh[0];
  __syncthreads();
  return r;
}

// op 0: sum a*b ; op 1: max |a|*(mask or 1) ; op 2: max |a|/(1+|b|)
__global__ __launch_bounds__(TPB) void k_red_partial(int op, int64_t n, const double* __restrict__ a, const double* __restrict__ b,
                                                     const double* __restrict__ mask, double* __restrict__ part) {
  __shared__ double sh[TPB];
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    if (op == 0) acc += a[i] * b[i];
    else if (op == 1) { double v = fabs(a[i]) * (mask ? mask[i] : 1.0); acc = (v > acc || v != v) ? v : acc; }
    else { double v = fabs(a[i]) / (1.0 + fabs(b[i])); acc = (v > acc || v != v) ? v : acc; }
  }
  double r = op == 0 ? block_sum(acc, sh) : block_max(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = r;
}

// reduces `np_` partials into scal[slot]; mode 0 sum, 1 max (NaN-propagating)
__global__ __launch_bounds__(TPB) void k_red_final(int mode, int npart, const double* __restrict__ part, double* scal, int slot) {
  __shared__ double sh[TPB];
  double acc = 0.0;
  for (int i = threadIdx.x; i < npart; i += blockDim.x) {
    double v = part[i];
    if (mode == 0) acc += v;
    else acc = (v > acc || v != v) ? v : acc;
  }
  double r = mode == 0 ? block_sum(acc, sh) : block_max(acc, sh);
  if (threadIdx.x == 0) scal[slot] = r;
}

// y = A x (node-block CSR, block size BS); masked: y *= free
template <int BS, typename VT>
__global__ __launch_bounds__(TPB) void k_spmv(int64_t np, const int32_t* __restrict__ rp, const int32_t* __restrict__ cols,
                                              const VT* __restrict__ vals, const double* __restrict__ x, double* __restrict__ y,
                                              const double* __restrict__ mask) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= np * BS) return;
  const int64_t a = t / BS;
  const int i = (int)(t - a * BS);
  double s = 0.0;
  for (int32_t k = rp[a]; k < rp[a + 1]; k++) {
    const int64_t b = cols[k];
    const VT* v = vals + (int64_t)k * BS * BS + i * BS;
#pragma unroll
    for (int j = 0; j < BS; j++) s += (double)v[j] * x[b * BS + j];
  }
  y[t] = mask ? s * mask[t] : s;
}

// block-Jacobi inverse of the diagonal node blocks, identity on prescribed rows/columns
template <int BS>
__global__ __launch_bounds__(TPB) void k_bjacobi(int64_t np, const int32_t* __restrict__ rp, const int32_t* __restrict__ cols,
                                                 const double* __restrict__ vals, const double* __restrict__ free_, double* __restrict__ binv) {
  const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= np) return;
  double A[BS][BS];
  int32_t k = rp[a];
  while (k < rp[a + 1] && cols[k] != a) k++;
  for (int i = 0; i < BS; i++)
    for (int j = 0; j < BS; j++) {
      bool fi = free_[a * BS + i] != 0.0, fj = free_[a * BS + j] != 0.0;
      A[i][j] = (fi && fj) ? vals[(int64_t)k * BS * BS + i * BS + j] : (i == j ? 1.0 : 0.0);
    }
  double Ai[BS][BS];
  if constexpr (BS == 2) {
    double d = A[0][0] * A[1][1] - A[0][1] * A[1][0];
    Ai[0][0] = A[1][1] / d; Ai[0][1] = -A[0][1] / d; Ai[1][0] = -A[1][0] / d; Ai[1][1] = A[0][0] / d;
  } else {
    double c00 = A[1][1] * A[2][2] - A[1][2] * A[2][1];
    double c01 = A[1][2] * A[2][0] - A[1][0] * A[2][2];
    double c02 = A[1][0] * A[2][1] - A[1][1] * A[2][0];
    double d = A[0][0] * c00 + A[0][1] * c01 + A[0][2] * c02, id = 1.0 / d;
    Ai[0][0] = c00 * id; Ai[0][1] = (A[0][2] * A[2][1] - A[0][1] * A[2][2]) * id; Ai[0][2] = (A[0][1] * A[1][2] - A[0][2] * A[1][1]) * id;
    Ai[1][0] = c01 * id; Ai[1][1] = (A[0][0] * A[2][2] - A[0][2] * A[2][0]) * id; Ai[1][2] = (A[0][2] * A[1][0] - A[0][0] * A[1][2]) * id;
    Ai[2][0] = c02 * id; Ai[2][1] = (A[0][1] * A[2][0] - A[0][0] * A[2][1]) * id; Ai[2][2] = (A[0][0] * A[1][1] - A[0][1] * A[1][0]) * id;
  }
  for (int i = 0; i < BS; i++)
    for (int j = 0; j < BS; j++) binv[a * BS * BS + i * BS + j] = Ai[i][j];
}

// PCG start: r = b*free, x = 0, z = M^-1 r, p = z ; partials of r.z and b.b
template <int BS>
__global__ __launch_bounds__(TPB) void k_pcg_start(int64_t np, const double* __restrict__ b, const double* __restrict__ free_,
                                                   const double* __restrict__ binv, double* __restrict__ x, double* __restrict__ r,
                                                   double* __restrict__ z, double* __restrict__ p, double* __restrict__ part_rz,
                                                   double* __restrict__ part_bb) {
  __shared__ double sh[TPB];
  const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double rz = 0.0, bb = 0.0;
  if (a < np) {
    double rv[BS];
    for (int i = 0; i < BS; i++) { rv[i] = b[a * BS + i] * free_[a * BS + i]; r[a * BS + i] = rv[i]; x[a * BS + i] = 0.0; bb += rv[i] * rv[i]; }
    for (int i = 0; i < BS; i++) {
      double s = 0.0;
      for (int j = 0; j < BS; j++) s += binv[a * BS * BS + i * BS + j] * rv[j];
      s *= free_[a * BS + i];
      z[a * BS + i] = s; p[a * BS + i] = s; rz += rv[i] * s;
    }
  }
  double s1 = block_sum(rz, sh), s2 = block_sum(bb, sh);
  if (threadIdx.x == 0) { part_rz[blockIdx.x] = s1; part_bb[blockIdx.x] = s2; }
}

// x += alpha p ; r -= alpha q ; z = M^-1 r ; partials of r.z and r.r
template <int BS>
__global__ __launch_bounds__(TPB) void k_pcg_xr(int64_t np, const double* __restrict__ scal, const double* __restrict__ p,
                                                const double* __restrict__ q, const double* __restrict__ free_, const double* __restrict__ binv,
                                                double* __restrict__ x, double* __restrict__ r, double* __restrict__ z,
                                                double* __restrict__ part_rz, double* __restrict__ part_rr) {
  __shared__ double sh[TPB];
  const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const double alpha = scal[S_ALPHA];
  double rz = 0.0, rr = 0.0;
  if (a < np) {
    double rv[BS];
    for (int i = 0; i < BS; i++) {
      const int64_t k = a * BS + i;
      x[k] += alpha * p[k];
      rv[i] = r[k] - alpha * q[k];
      r[k] = rv[i];
      rr += rv[i] * rv[i];
    }
    for (int i = 0; i < BS; i++) {
      double s = 0.0;
      for (int j = 0; j < BS; j++) s += binv[a * BS * BS + i * BS + j] * rv[j];
      s *= free_[a * BS + i];
      z[a * BS + i] = s; rz += rv[i] * s;
    }
  }
  double s1 = block_sum(rz, sh), s2 = block_sum(rr, sh);
  if (threadIdx.x == 0) { part_rz[blockIdx.x] = s1; part_rr[blockIdx.x] = s2; }
}

// scalar updates: mode 0: alpha = rz/pq ; mode 1: beta = rznew/rz, rz = rznew ; mode 2: rz = rznew (start)
__global__ void k_pcg_scalar(int mode, double* scal) {
  if (threadIdx.x != 0) return;
  if (mode == 0) scal[S_ALPHA] = scal[S_PQ] != 0.0 ? scal[S_RZ] / scal[S_PQ] : 0.0;
  else if (mode == 1) { scal[S_BETA] = scal[S_RZ] != 0.0 ? scal[S_RZNEW] / scal[S_RZ] : 0.0; scal[S_RZ] = scal[S_RZNEW]; }
  else scal[S_RZ] = scal[S_RZNEW];
}

__global__ __launch_bounds__(TPB) void k_pcg_p(int64_t n, const double* __restrict__ scal, const double* __restrict__ z, double* __restrict__ p) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = z[i] + scal[S_BETA] * p[i];
}

// R = F_int - F_ext ; LMM rows (bc_prescribed.rs:99-108) or RSM (ignored rows = 0)
__global__ __launch_bounds__(TPB) void k_residual(int64_t n, const double* __restrict__ fint, const double* __restrict__ fext,
                                                  double* __restrict__ rr) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) rr[i] = fint[i] - fext[i];
}

__global__ __launch_bounds__(TPB) void k_prescribed_rows(int64_t np_, int lmm, const int64_t* __restrict__ peq, const double* __restrict__ lam,
                                                         const double* __restrict__ u, const double* __restrict__ cval,
                                                         double* __restrict__ rr, double* __restrict__ g) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np_) return;
  const int64_t eq = peq[p];
  if (lmm) { rr[eq] += lam[p]; g[p] = u[eq] - cval[p]; }
  else { rr[eq] = 0.0; g[p] = 0.0; }
}

// t = 0 except t[eq_p] = g_p
__global__ __launch_bounds__(TPB) void k_scatter_g(int64_t np_, const int64_t* __restrict__ peq, const double* __restrict__ g, double* __restrict__ t) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < np_) t[peq[p]] = g[p];
}

// b = (rr - q) * free
__global__ __launch_bounds__(TPB) void k_rhs(int64_t n, const double* __restrict__ rr, const double* __restrict__ q,
                                             const double* __restrict__ free_, double* __restrict__ b) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (rr[i] - q[i]) * free_[i];
}

// y_p = rr[eq] - (K x)[eq]
__global__ __launch_bounds__(TPB) void k_mult_corr(int64_t np_, const int64_t* __restrict__ peq, const double* __restrict__ rr,
                                                   const double* __restrict__ kx, double* __restrict__ y) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < np_) y[p] = rr[peq[p]] - kx[peq[p]];
}

// u -= x ; ddu -= x   (solver_data.rs:135-152)
__global__ __launch_bounds__(TPB) void k_update_u(int64_t n, const double* __restrict__ x, double* __restrict__ u, double* __restrict__ ddu) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) { u[i] -= x[i]; ddu[i] -= x[i]; }
}

__global__ __launch_bounds__(TPB) void k_axpy(int64_t n, double a, const double* __restrict__ x, double* __restrict__ y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] += a * x[i];
}

inline unsigned nblk(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + TPB - 1) / TPB); }

}  // namespace

// ================================================================================================
double* scal_ptr(pm_fem* f, int slot) { return f->d_scal + slot; }

static void reduce(pm_fem* f, int op, int64_t n, const double* a, const double* b, const double* mask, int slot) {
  unsigned grid = (unsigned)std::min<int64_t>(RED_BLOCKS, std::max<int64_t>(1, (n + TPB - 1) / TPB));
  hipLaunchKernelGGL(k_red_partial, dim3(grid), dim3(TPB), 0, f->stream, op, n, a, b, mask, f->d_red);
  hipLaunchKernelGGL(k_red_final, dim3(1), dim3(TPB), 0, f->stream, op == 0 ? 0 : 1, (int)grid, f->d_red, f->d_scal, slot);
}

static double read_scal(pm_fem* f, int slot) {
  double v = 0.0;
  (void)hipMemcpyAsync(&v, f->d_scal + slot, sizeof(double), hipMemcpyDeviceToHost, f->stream);
  (void)hipStreamSynchronize(f->stream);
  return v;
}

double dev_max_abs(pm_fem* f, const double* x, int64_t n, const double* mask) {
  if (n == 0) return 0.0;
  reduce(f, 1, n, x, nullptr, mask, S_MAX);
  return read_scal(f, S_MAX);
}

double dev_max_rel(pm_fem* f, const double* x, const double* x0, int64_t n) {
  if (n == 0) return 0.0;
  reduce(f, 2, n, x, x0, nullptr, S_MAX);
  return read_scal(f, S_MAX);
}

double dev_dot(pm_fem* f, const double* a, const double* b, int64_t n) {
  reduce(f, 0, n, a, b, nullptr, S_TMP);
  return read_scal(f, S_TMP);
}

void dev_spmv(pm_fem* f, const double* x, double* y, bool masked) {
  const int64_t n = f->npoint * f->bs;
  const double* mask = masked ? f->d_free : nullptr;
  const bool lowp = f->cfg.tangent_precision != 0 && f->d_kval32;
  if (f->bs == 2) {
    if (lowp) hipLaunchKernelGGL((k_spmv<2, float>), dim3(nblk(n)), dim3(TPB), 0, f->stream, f->npoint, f->d_browptr, f->d_bcols, f->d_kval32, x, y, mask);
    else hipLaunchKernelGGL((k_spmv<2, double>), dim3(nblk(n)), dim3(TPB), 0, f->stream, f->npoint, f->d_browptr, f->d_bcols, f->d_kval, x, y, mask);
  } else {
    if (lowp) hipLaunchKernelGGL((k_spmv<3, float>), dim3(nblk(n)), dim3(TPB), 0, f->stream, f->npoint, f->d_browptr, f->d_bcols, f->d_kval32, x, y, mask);
    else hipLaunchKernelGGL((k_spmv<3, double>), dim3(nblk(n)), dim3(TPB), 0, f->stream, f->npoint, f->d_browptr, f->d_bcols, f->d_kval, x, y, mask);
  }
}

void build_block_jacobi(pm_fem* f) {
  if (f->bs == 2) hipLaunchKernelGGL((k_bjacobi<2>), dim3(nblk(f->npoint)), dim3(TPB), 0, f->stream, f->npoint, f->d_browptr, f->d_bcols, f->d_kval, f->d_free, f->d_binv);
  else hipLaunchKernelGGL((k_bjacobi<3>), dim3(nblk(f->npoint)), dim3(TPB), 0, f->stream, f->npoint, f->d_browptr, f->d_bcols, f->d_kval, f->d_free, f->d_binv);
}

// Preconditioned CG on the free DOFs of K (prescribed entries of b ignored, of x zero).
int pcg_solve(pm_fem* f, const double* d_b, double* d_x, int* iters) {
  const int64_t np = f->npoint, n = np * f->bs;
  double* r = f->d_pcg;
  double* z = r + n;
  double* p = z + n;
  double* q = p + n;
  const unsigned gn = nblk(np);
  double* part_a = f->d_red;
  double* part_b = f->d_red + gn;
  if (f->bs == 2) hipLaunchKernelGGL((k_pcg_start<2>), dim3(gn), dim3(TPB), 0, f->stream, np, d_b, f->d_free, f->d_binv, d_x, r, z, p, part_a, part_b);
  else hipLaunchKernelGGL((k_pcg_start<3>), dim3(gn), dim3(TPB), 0, f->stream, np, d_b, f->d_free, f->d_binv, d_x, r, z, p, part_a, part_b);
  hipLaunchKernelGGL(k_red_final, dim3(1), dim3(TPB), 0, f->stream, 0, (int)gn, part_a, f->d_scal, (int)S_RZNEW);
  hipLaunchKernelGGL(k_red_final, dim3(1), dim3(TPB), 0, f->stream, 0, (int)gn, part_b, f->d_scal, (int)S_BB);
  hipLaunchKernelGGL(k_pcg_scalar, dim3(1), dim3(64), 0, f->stream, 2, f->d_scal);
  double bb = read_scal(f, S_BB);
  *iters = 0;
  if (bb == 0.0) return 0;
  if (!std::isfinite(bb)) { f->err = "Found NaN or Inf"; return 1; }
  const double tol2 = f->cfg.pcg_rtol * f->cfg.pcg_rtol * bb;
  const int check_every = 8;
  int it = 0;
  double rr = bb;
  while (it < f->cfg.pcg_max_iter) {
    dev_spmv(f, p, q, true);
    reduce(f, 0, n, p, q, nullptr, S_PQ);
    hipLaunchKernelGGL(k_pcg_scalar, dim3(1), dim3(64), 0, f->stream, 0, f->d_scal);
    if (f->bs == 2) hipLaunchKernelGGL((k_pcg_xr<2>), dim3(gn), dim3(TPB), 0, f->stream, np, f->d_scal, p, q, f->d_free, f->d_binv, d_x, r, z, part_a, part_b);
    else hipLaunchKernelGGL((k_pcg_xr<3>), dim3(gn), dim3(TPB), 0, f->stream, np, f->d_scal, p, q, f->d_free, f->d_binv, d_x, r, z, part_a, part_b);
    hipLaunchKernelGGL(k_red_final, dim3(1), dim3(TPB), 0, f->stream, 0, (int)gn, part_a, f->d_scal, (int)S_RZNEW);
    hipLaunchKernelGGL(k_red_final, dim3(1), dim3(TPB), 0, f->stream, 0, (int)gn, part_b, f->d_scal, (int)S_RR);
    hipLaunchKernelGGL(k_pcg_scalar, dim3(1), dim3(64), 0, f->stream, 1, f->d_scal);
    hipLaunchKernelGGL(k_pcg_p, dim3(nblk(n)), dim3(TPB), 0, f->stream, n, f->d_scal, z, p);
    it++;
    if (it % check_every == 0 || it == f->cfg.pcg_max_iter) {
      rr = read_scal(f, S_RR);
      if (!std::isfinite(rr)) { f->err = "Found NaN or Inf in mdu"; *iters = it; return 1; }
      if (rr <= tol2) break;
    }
  }
  *iters = it;
  f->stats.n_pcg_iterations += it;
  if (rr > tol2) { f->err = "PCG did not converge"; return 2; }
  return 0;
}

// ---- Newton vector steps ------------------------------------------------------------------------
void dev_residual(pm_fem* f, double* g) {
  hipLaunchKernelGGL(k_residual, dim3(nblk(f->ndof)), dim3(TPB), 0, f->stream, f->ndof, f->d_fint, f->d_fext, f->d_rr);
  if (f->nprescribed)
    hipLaunchKernelGGL(k_prescribed_rows, dim3(nblk(f->nprescribed)), dim3(TPB), 0, f->stream, f->nprescribed,
                       f->cfg.lagrange_mult_method, f->d_peq, f->d_lam, f->d_u, f->d_cval, f->d_rr, g);
}

// b = (rr - K t) * free with t = g on prescribed DOFs
void dev_rhs(pm_fem* f, const double* g, double* t, double* kt, double* b) {
  dev_fill(f, t, 0.0, f->ndof);
  if (f->nprescribed) hipLaunchKernelGGL(k_scatter_g, dim3(nblk(f->nprescribed)), dim3(TPB), 0, f->stream, f->nprescribed, f->d_peq, g, t);
  dev_spmv(f, t, kt, false);
  hipLaunchKernelGGL(k_rhs, dim3(nblk(f->ndof)), dim3(TPB), 0, f->stream, f->ndof, f->d_rr, kt, f->d_free, b);
}

// x[eq_p] = g_p ; y_p = rr[eq_p] - (K x)[eq_p]
void dev_complete_correction(pm_fem* f, const double* g, double* x, double* kx, double* y) {
  if (!f->nprescribed) return;
  hipLaunchKernelGGL(k_scatter_g, dim3(nblk(f->nprescribed)), dim3(TPB), 0, f->stream, f->nprescribed, f->d_peq, g, x);
  dev_spmv(f, x, kx, false);
  hipLaunchKernelGGL(k_mult_corr, dim3(nblk(f->nprescribed)), dim3(TPB), 0, f->stream, f->nprescribed, f->d_peq, f->d_rr, kx, y);
}

void dev_update_primary(pm_fem* f, const double* x, const double* y) {
  hipLaunchKernelGGL(k_update_u, dim3(nblk(f->ndof)), dim3(TPB), 0, f->stream, f->ndof, x, f->d_u, f->d_ddu);
  if (f->nprescribed && f->cfg.lagrange_mult_method)
    hipLaunchKernelGGL(k_axpy, dim3(nblk(f->nprescribed)), dim3(TPB), 0, f->stream, f->nprescribed, -1.0, y, f->d_lam);
}

}  // namespace pm
