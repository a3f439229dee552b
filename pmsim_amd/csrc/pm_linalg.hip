// Device linear algebra of the Newton driver: node-block CSR SpMV, block-Jacobi PCG (FP64 vectors,
// FP64 or FP32 matrix values), deterministic fixed-shape reductions, and the vector updates of
// pmsim's residual / multiplier / primary-variable steps (solver_data.rs:117-152, bc_prescribed.rs:99-108).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "pm_internal.hpp"
#include "pm_spmv.hpp"

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

// y = A x (node blocks of size BS, upper triangle stored); masked: y *= free
template <int BS, typename VT>
__global__ __launch_bounds__(TPB) void k_spmv(int64_t np, KSym<VT> K, const double* __restrict__ x, double* __restrict__ y,
                                              const double* __restrict__ mask) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= np * BS) return;
  const int64_t a = t / BS;
  const int i = (int)(t - a * BS);
  const double s = ksym_row<BS, VT>(K, a, i, x);
  y[t] = mask ? s * mask[t] : s;
}

// block-Jacobi inverse of the diagonal node blocks (from the K values in the tangent's precision),
// identity on prescribed rows/columns
template <int BS, typename KT>
__global__ __launch_bounds__(TPB) void k_bjacobi(int64_t np, const int32_t* __restrict__ urp, const KT* __restrict__ vals,
                                                 const double* __restrict__ free_, double* __restrict__ binv) {
  const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= np) return;
  double A[BS][BS];
  const int32_t k = urp[a];  // the diagonal block is the first stored block of its row
  for (int i = 0; i < BS; i++)
    for (int j = 0; j < BS; j++) {
      bool fi = free_[a * BS + i] != 0.0, fj = free_[a * BS + j] != 0.0;
      A[i][j] = (fi && fj) ? (double)vals[(int64_t)k * BS * BS + i * BS + j] : (i == j ? 1.0 : 0.0);
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
// R = F_int - ell F_ext (solver_data.rs:117-120); ell = 1 gives the bits of F_int - F_ext
__global__ __launch_bounds__(TPB) void k_residual(int64_t n, const double* __restrict__ fint, const double* __restrict__ fext,
                                                  double ell, double* __restrict__ rr) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) rr[i] = fint[i] - ell * fext[i];
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

// arc-length predictor (arc_length_control.rs:69-79): u = (1 + alpha) u_old - alpha u_anc (alpha = 0 keeps
// u), du = u - u_old
__global__ __launch_bounds__(TPB) void k_arc_trial(int64_t n, int extrapolate, double alpha, const double* __restrict__ u_old,
                                                   const double* __restrict__ u_anc, double* __restrict__ u, double* __restrict__ du) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (extrapolate) u[i] = (1.0 + alpha) * u_old[i] - alpha * u_anc[i];
  du[i] = u[i] - u_old[i];
}

// out = x * free
__global__ __launch_bounds__(TPB) void k_masked(int64_t n, const double* __restrict__ x, const double* __restrict__ free_,
                                                double* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = x[i] * free_[i];
}

// ---- fused PCG iteration (3 kernels, fixed grid; every block reduces the previous kernel's partials
//      in the same order, so the scalars are identical in all blocks and across runs) ----------------
constexpr int PCG_TPB = 256;

__device__ __forceinline__ double sum_partials(const double* __restrict__ part, int n, double* sh) {
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) acc += part[i];
  return block_sum(acc, sh);
}

template <int BS>
__device__ __forceinline__ void bjacobi_apply(int64_t a, const double* __restrict__ binv, const double* __restrict__ free_,
                                              const double* rv, double* zv) {
#pragma unroll
  for (int i = 0; i < BS; i++) {
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < BS; j++) s += binv[a * BS * BS + i * BS + j] * rv[j];
    zv[i] = s * free_[a * BS + i];
  }
}

// full-pattern K from the stored blocks: block k of the full pattern is the stored block f2u[k], read
// transposed when the sign bit is set (the lower blocks)
template <int BS, typename VT>
__global__ __launch_bounds__(TPB) void k_widen(int64_t nnzb, const int32_t* __restrict__ f2u, const VT* __restrict__ up,
                                               VT* __restrict__ full) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nnzb) return;
  const uint32_t m = (uint32_t)f2u[k];
  const VT* s = up + (int64_t)(m & 0x7fffffffu) * BS * BS;
  VT v[BS * BS];
#pragma unroll
  for (int q = 0; q < BS * BS; q++) v[q] = s[q];
  VT* d = full + k * BS * BS;
  const bool tr = (m & 0x80000000u) != 0;
#pragma unroll
  for (int i = 0; i < BS; i++)
#pragma unroll
    for (int j = 0; j < BS; j++) d[i * BS + j] = tr ? v[j * BS + i] : v[i * BS + j];
}

// K1: q = (A p) * free over the full-pattern copy of K, LPR lanes per block row (full_row_lanes);
// partial p.q per workgroup
template <int BS, typename VT, int LPR>
__global__ __launch_bounds__(PCG_TPB) void k_pcg_q_rows(int64_t np, const int32_t* __restrict__ rp, const int32_t* __restrict__ cols,
                                                        const VT* __restrict__ vals, const double* __restrict__ free_,
                                                        const double* __restrict__ p, double* __restrict__ q, double* __restrict__ part) {
  static_assert(LPR >= 1 && LPR <= 64 && (LPR & (LPR - 1)) == 0, "lanes per row: a power of two within a wave");
  __shared__ double sh[PCG_TPB];
  const int lane = (int)(threadIdx.x & (LPR - 1));
  const int64_t ngroups = (int64_t)gridDim.x * blockDim.x / LPR;
  double acc = 0.0;
  // a depends on the group only: the lanes of a group run the same rows and reach the shuffles together.
  // 3x3 blocks: two rows per pass (measured 7 % faster on Hex20; slower on 2D meshes); the p.q sum takes
  // the rows in the same order as one row per pass would, so the bits are the same either way.
  auto finish = [&](int64_t a, const double* y) {
    if (lane == 0) {
#pragma unroll
      for (int i = 0; i < BS; i++) {
        const int64_t t = a * BS + i;
        const double sv = y[i] * free_[t];
        q[t] = sv;
        acc += p[t] * sv;
      }
    }
  };
  int64_t a = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / LPR;
  if constexpr (BS == 3) {
    for (; a + ngroups < np; a += 2 * ngroups) {
      double y0[BS], y1[BS];
      full_row_lanes2<BS, VT>(rp, cols, vals, a, a + ngroups, lane, LPR, p, y0, y1);
      finish(a, y0);
      finish(a + ngroups, y1);
    }
  }
  for (; a < np; a += ngroups) {
    double y[BS];
    full_row_lanes<BS, VT>(rp, cols, vals, a, lane, LPR, p, y);
    finish(a, y);
  }
  const double r = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = r;
}

// K2: alpha = rz / pq ; x += alpha p ; r -= alpha q ; z = M^-1 r ; partials r.z, r.r
template <int BS>
__global__ __launch_bounds__(PCG_TPB) void k_pcg_xrz(int64_t np, int npart, const double* __restrict__ part_pq, const double* __restrict__ rz_in,
                                                     const double* __restrict__ p, const double* __restrict__ q,
                                                     const double* __restrict__ free_, const double* __restrict__ binv,
                                                     double* __restrict__ x, double* __restrict__ r, double* __restrict__ z,
                                                     double* __restrict__ part_rz, double* __restrict__ part_rr) {
  __shared__ double sh[PCG_TPB];
  const double pq = sum_partials(part_pq, npart, sh);
  const double alpha = pq != 0.0 ? (*rz_in) / pq : 0.0;
  double rz = 0.0, rr = 0.0;
  for (int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; a < np; a += (int64_t)gridDim.x * blockDim.x) {
    double rv[BS], zv[BS];
#pragma unroll
    for (int i = 0; i < BS; i++) {
      const int64_t k = a * BS + i;
      x[k] += alpha * p[k];
      rv[i] = r[k] - alpha * q[k];
      r[k] = rv[i];
      rr += rv[i] * rv[i];
    }
    bjacobi_apply<BS>(a, binv, free_, rv, zv);
#pragma unroll
    for (int i = 0; i < BS; i++) { z[a * BS + i] = zv[i]; rz += rv[i] * zv[i]; }
  }
  const double s1 = block_sum(rz, sh), s2 = block_sum(rr, sh);
  if (threadIdx.x == 0) { part_rz[blockIdx.x] = s1; part_rr[blockIdx.x] = s2; }
}

// K3: rz' = sum ; beta = rz'/rz ; p = z + beta p ; block 0 stores rz' (other slot) and r.r
__global__ __launch_bounds__(PCG_TPB) void k_pcg_p2(int64_t n, int npart, const double* __restrict__ part_rz, const double* __restrict__ part_rr,
                                                    const double* __restrict__ rz_old, double* __restrict__ rz_new_out,
                                                    double* __restrict__ rr_out, const double* __restrict__ z, double* __restrict__ p) {
  __shared__ double sh[PCG_TPB];
  const double rzn = sum_partials(part_rz, npart, sh);
  const double beta = (*rz_old) != 0.0 ? rzn / (*rz_old) : 0.0;
  if (blockIdx.x == 0) {
    const double rr = sum_partials(part_rr, npart, sh);
    if (threadIdx.x == 0) { *rz_new_out = rzn; *rr_out = rr; }
  }
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = z[i] + beta * p[i];
}

// start: r = b*free, x = 0, z = M^-1 r, p = z ; partials r.z and b.b
template <int BS>
__global__ __launch_bounds__(PCG_TPB) void k_pcg_init(int64_t np, const double* __restrict__ b, const double* __restrict__ free_,
                                                      const double* __restrict__ binv, double* __restrict__ x, double* __restrict__ r,
                                                      double* __restrict__ z, double* __restrict__ p, double* __restrict__ part_rz,
                                                      double* __restrict__ part_bb) {
  __shared__ double sh[PCG_TPB];
  double rz = 0.0, bb = 0.0;
  for (int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; a < np; a += (int64_t)gridDim.x * blockDim.x) {
    double rv[BS], zv[BS];
#pragma unroll
    for (int i = 0; i < BS; i++) { rv[i] = b[a * BS + i] * free_[a * BS + i]; r[a * BS + i] = rv[i]; x[a * BS + i] = 0.0; bb += rv[i] * rv[i]; }
    bjacobi_apply<BS>(a, binv, free_, rv, zv);
#pragma unroll
    for (int i = 0; i < BS; i++) { z[a * BS + i] = zv[i]; p[a * BS + i] = zv[i]; rz += rv[i] * zv[i]; }
  }
  const double s1 = block_sum(rz, sh), s2 = block_sum(bb, sh);
  if (threadIdx.x == 0) { part_rz[blockIdx.x] = s1; part_bb[blockIdx.x] = s2; }
}

__global__ void k_pcg_init_fin(int npart, const double* __restrict__ part_rz, const double* __restrict__ part_bb, double* rz_out, double* bb_out) {
  __shared__ double sh[PCG_TPB];
  const double a = sum_partials(part_rz, npart, sh);
  const double b = sum_partials(part_bb, npart, sh);
  if (threadIdx.x == 0) { *rz_out = a; *bb_out = b; }
}

// K1 over the stored upper blocks (the lower ones read through their mirrors; no full-pattern copy):
// q = (A p) * free, one thread per row component; partial p.q per workgroup
template <int BS, typename VT>
__global__ __launch_bounds__(PCG_TPB) void k_pcg_q_sym(int64_t np, KSym<VT> K, const double* __restrict__ free_,
                                                       const double* __restrict__ p, double* __restrict__ q, double* __restrict__ part) {
  __shared__ double sh[PCG_TPB];
  double acc = 0.0;
  const int64_t n = np * BS;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t a = t / BS;
    const int i = (int)(t - a * BS);
    const double sv = ksym_row<BS, VT>(K, a, i, p) * free_[t];
    q[t] = sv;
    acc += p[t] * sv;
  }
  const double r = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = r;
}

// ---- single-workgroup persistent PCG for small systems: the whole loop runs in one launch ----------
constexpr int SMALL_TPB = 1024;

__device__ __forceinline__ void block_sum2(double& a, double& b, double* sh) {
  const int t = threadIdx.x;
  sh[t] = a; sh[SMALL_TPB + t] = b;
  __syncthreads();
  for (int s = SMALL_TPB / 2; s > 0; s >>= 1) {
    if (t < s) { sh[t] += sh[t + s]; sh[SMALL_TPB + t] += sh[SMALL_TPB + t + s]; }
    __syncthreads();
  }
  a = sh[0]; b = sh[SMALL_TPB];
  __syncthreads();
}

template <int BS, typename VT>
__global__ __launch_bounds__(SMALL_TPB) void k_pcg_small(int64_t np, KSym<VT> K, const double* __restrict__ free_,
                                                         const double* __restrict__ binv, const double* __restrict__ b,
                                                         double* __restrict__ x, double* __restrict__ r, double* __restrict__ z,
                                                         double* __restrict__ p, double* __restrict__ q, double rtol2, int maxit,
                                                         double* __restrict__ out /* [iterations, rr, bb] */) {
  __shared__ double sh[2 * SMALL_TPB];
  const int64_t n = np * BS;
  double rz = 0.0, bb = 0.0;
  for (int64_t a = threadIdx.x; a < np; a += SMALL_TPB) {
    double rv[BS], zv[BS];
#pragma unroll
    for (int i = 0; i < BS; i++) { rv[i] = b[a * BS + i] * free_[a * BS + i]; r[a * BS + i] = rv[i]; x[a * BS + i] = 0.0; bb += rv[i] * rv[i]; }
    bjacobi_apply<BS>(a, binv, free_, rv, zv);
#pragma unroll
    for (int i = 0; i < BS; i++) { z[a * BS + i] = zv[i]; p[a * BS + i] = zv[i]; rz += rv[i] * zv[i]; }
  }
  block_sum2(rz, bb, sh);
  const double tol2 = rtol2 * bb;
  int it = 0;
  double rr = bb;
  if (bb > 0.0) {
    while (it < maxit) {
      __syncthreads();  // p complete
      double pq = 0.0, dummy = 0.0;
      for (int64_t t = threadIdx.x; t < n; t += SMALL_TPB) {
        const int64_t a = t / BS;
        const int i = (int)(t - a * BS);
        double s = ksym_row<BS, VT>(K, a, i, p);
        s *= free_[t];
        q[t] = s;
        pq += p[t] * s;
      }
      block_sum2(pq, dummy, sh);
      const double alpha = pq != 0.0 ? rz / pq : 0.0;
      double rzn = 0.0;
      rr = 0.0;
      for (int64_t a = threadIdx.x; a < np; a += SMALL_TPB) {
        double rv[BS], zv[BS];
#pragma unroll
        for (int i = 0; i < BS; i++) {
          const int64_t k = a * BS + i;
          x[k] += alpha * p[k];
          rv[i] = r[k] - alpha * q[k];
          r[k] = rv[i];
          rr += rv[i] * rv[i];
        }
        bjacobi_apply<BS>(a, binv, free_, rv, zv);
#pragma unroll
        for (int i = 0; i < BS; i++) { z[a * BS + i] = zv[i]; rzn += rv[i] * zv[i]; }
      }
      block_sum2(rzn, rr, sh);
      it++;
      if (rr <= tol2 || !(rr == rr)) break;
      const double beta = rz != 0.0 ? rzn / rz : 0.0;
      rz = rzn;
      for (int64_t i = threadIdx.x; i < n; i += SMALL_TPB) p[i] = z[i] + beta * p[i];
    }
  }
  if (threadIdx.x == 0) { out[0] = (double)it; out[1] = rr; out[2] = bb; }
}

inline unsigned nblk(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + TPB - 1) / TPB); }

}  // namespace

// ================================================================================================
double* scal_ptr(pm_fem* f, int slot) { return f->d_scal + slot; }

template <typename VT>
static KSym<VT> ksym(const pm_fem* f, const VT* vals) { return KSym<VT>{f->d_urp, f->d_ucol, f->d_lrp, f->d_lcol, f->d_lidx, vals}; }

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
    if (lowp) hipLaunchKernelGGL((k_spmv<2, float>), dim3(nblk(n)), dim3(TPB), 0, f->stream, f->npoint, ksym(f, f->d_kval32), x, y, mask);
    else hipLaunchKernelGGL((k_spmv<2, double>), dim3(nblk(n)), dim3(TPB), 0, f->stream, f->npoint, ksym(f, f->d_kval), x, y, mask);
  } else {
    if (lowp) hipLaunchKernelGGL((k_spmv<3, float>), dim3(nblk(n)), dim3(TPB), 0, f->stream, f->npoint, ksym(f, f->d_kval32), x, y, mask);
    else hipLaunchKernelGGL((k_spmv<3, double>), dim3(nblk(n)), dim3(TPB), 0, f->stream, f->npoint, ksym(f, f->d_kval), x, y, mask);
  }
}

void build_block_jacobi(pm_fem* f) {
  const unsigned g = nblk(f->npoint);
  if (f->d_kval) {
    if (f->bs == 2) hipLaunchKernelGGL((k_bjacobi<2, double>), dim3(g), dim3(TPB), 0, f->stream, f->npoint, f->d_urp, f->d_kval, f->d_free, f->d_binv);
    else hipLaunchKernelGGL((k_bjacobi<3, double>), dim3(g), dim3(TPB), 0, f->stream, f->npoint, f->d_urp, f->d_kval, f->d_free, f->d_binv);
  } else {
    if (f->bs == 2) hipLaunchKernelGGL((k_bjacobi<2, float>), dim3(g), dim3(TPB), 0, f->stream, f->npoint, f->d_urp, f->d_kval32, f->d_free, f->d_binv);
    else hipLaunchKernelGGL((k_bjacobi<3, float>), dim3(g), dim3(TPB), 0, f->stream, f->npoint, f->d_urp, f->d_kval32, f->d_free, f->d_binv);
  }
}

// Preconditioned CG on the free DOFs of K (prescribed entries of b ignored, of x zero).
// Small systems run the whole loop in one single-workgroup launch; larger ones run three fused kernels
// per iteration on a fixed grid and check convergence on the host every few iterations.
int pcg_solve(pm_fem* f, const double* d_b, double* d_x, int* iters) {
  const int64_t np = f->npoint, n = np * f->bs;
  double* r = f->d_pcg;
  double* z = r + n;
  double* p = z + n;
  double* q = p + n;
  const bool lowp = f->cfg.tangent_precision != 0 && f->d_kval32;
  const double rtol2 = f->cfg.pcg_rtol * f->cfg.pcg_rtol;
  *iters = 0;
  if (n <= f->pcg_small_max) {
    double* out = f->d_scal + 20;
#define PM_SMALL(BS, VT, VALS) hipLaunchKernelGGL((k_pcg_small<BS, VT>), dim3(1), dim3(SMALL_TPB), 0, f->stream, np, ksym(f, VALS), \
                                                  f->d_free, f->d_binv, d_b, d_x, r, z, p, q, rtol2, f->cfg.pcg_max_iter, out)
    if (f->bs == 2) { if (lowp) PM_SMALL(2, float, f->d_kval32); else PM_SMALL(2, double, f->d_kval); }
    else { if (lowp) PM_SMALL(3, float, f->d_kval32); else PM_SMALL(3, double, f->d_kval); }
#undef PM_SMALL
    double h[3];
    PM_HIP(hipMemcpyAsync(h, out, sizeof(h), hipMemcpyDeviceToHost, f->stream));
    PM_HIP(hipStreamSynchronize(f->stream));
    *iters = (int)h[0];
    f->stats.n_pcg_iterations += *iters;
    if (!std::isfinite(h[1]) || !std::isfinite(h[2])) { f->err = "Found NaN or Inf in mdu"; return 1; }
    if (h[2] == 0.0 || h[1] <= rtol2 * h[2]) return 0;
    f->err = "PCG did not converge";
    return 2;
  }
  // K widened to the full pattern once per solve; the iterations stream rows of it (same values and
  // column order as the stored blocks give)
  if (full_pattern_alloc(f)) return 3;  // out of device memory: the caller may fall back to the host solver
  {
    const unsigned gw = (unsigned)((f->nnzb + TPB - 1) / TPB);
#define PM_WIDEN(BS) do { if (lowp) hipLaunchKernelGGL((k_widen<BS, float>), dim3(gw), dim3(TPB), 0, f->stream, f->nnzb, f->d_f2u, f->d_kval32, f->d_kfull32); \
                          else hipLaunchKernelGGL((k_widen<BS, double>), dim3(gw), dim3(TPB), 0, f->stream, f->nnzb, f->d_f2u, f->d_kval, f->d_kfull); } while (0)
    if (f->bs == 2) PM_WIDEN(2);
    else PM_WIDEN(3);
#undef PM_WIDEN
  }
  if (f->engine) return engine_solve(f->engine, f, d_b, d_x, iters);  // batched with the other served requests
  // grids: rows for the SpMV (rbq), nodes for the vector updates (rb); partial arrays sized accordingly
  const int rb = (int)std::min<int64_t>(2048, std::max<int64_t>(1, (np + PCG_TPB - 1) / PCG_TPB));
  const int lpr = lanes_per_row(f->nnzb, np);
  const int rbq = spmv_workgroups(np, lpr, PCG_TPB);
  double* pa = f->d_red;            // partials: pa (rbq), pb (rb), pc (rb)
  double* pb = pa + rbq;
  double* pc = pb + rb;
  double* rzs = f->d_scal + 16;      // rz double buffer [16], [17]; rr at [18]; bb at [19]
  double* rr_s = f->d_scal + 18;
  double* bb_s = f->d_scal + 19;
  if (f->bs == 2) hipLaunchKernelGGL((k_pcg_init<2>), dim3(rb), dim3(PCG_TPB), 0, f->stream, np, d_b, f->d_free, f->d_binv, d_x, r, z, p, pa, pb);
  else hipLaunchKernelGGL((k_pcg_init<3>), dim3(rb), dim3(PCG_TPB), 0, f->stream, np, d_b, f->d_free, f->d_binv, d_x, r, z, p, pa, pb);
  hipLaunchKernelGGL(k_pcg_init_fin, dim3(1), dim3(PCG_TPB), 0, f->stream, rb, pa, pb, rzs, bb_s);
  double bb = 0.0;
  PM_HIP(hipMemcpyAsync(&bb, bb_s, sizeof(double), hipMemcpyDeviceToHost, f->stream));
  PM_HIP(hipStreamSynchronize(f->stream));
  if (bb == 0.0) return 0;
  if (!std::isfinite(bb)) { f->err = "Found NaN or Inf"; return 1; }
  const double tol2 = rtol2 * bb;
  const int check_every = 16;  // even, so a chunk starts on the same rz slot every time
  auto launch_q_rows = [&](int grid) {
#define PM_QROWS(BS, L) do { if (lowp) hipLaunchKernelGGL((k_pcg_q_rows<BS, float, L>), dim3(grid), dim3(PCG_TPB), 0, f->stream, np, f->d_frp, f->d_fcol, f->d_kfull32, f->d_free, p, q, pa); \
                             else hipLaunchKernelGGL((k_pcg_q_rows<BS, double, L>), dim3(grid), dim3(PCG_TPB), 0, f->stream, np, f->d_frp, f->d_fcol, f->d_kfull, f->d_free, p, q, pa); } while (0)
#define PM_QROWS_L(BS) do { if (lpr == 32) PM_QROWS(BS, 32); else PM_QROWS(BS, 8); } while (0)
    if (f->bs == 2) PM_QROWS_L(2);
    else PM_QROWS_L(3);
#undef PM_QROWS_L
#undef PM_QROWS
  };
  // one PCG iteration = 3 launches; iteration `it` reads rz from slot it&1 and writes slot (it+1)&1
  auto iterate = [&](int it) {
    const double* rz_old = rzs + (it & 1);
    double* rz_new = rzs + ((it + 1) & 1);
    launch_q_rows(rbq);
    if (f->bs == 2) hipLaunchKernelGGL((k_pcg_xrz<2>), dim3(rb), dim3(PCG_TPB), 0, f->stream, np, rbq, pa, rz_old, p, q, f->d_free, f->d_binv, d_x, r, z, pb, pc);
    else hipLaunchKernelGGL((k_pcg_xrz<3>), dim3(rb), dim3(PCG_TPB), 0, f->stream, np, rbq, pa, rz_old, p, q, f->d_free, f->d_binv, d_x, r, z, pb, pc);
    hipLaunchKernelGGL(k_pcg_p2, dim3(rb), dim3(PCG_TPB), 0, f->stream, n, rb, pb, pc, rz_old, rz_new, rr_s, z, p);
  };
  // a whole chunk of iterations between two convergence checks is captured once per handle (and
  // solution buffer) and replayed as one graph launch
  if (f->cfg.use_graphs && f->pcg_graph_x != d_x && f->pcg_graph) { (void)hipGraphExecDestroy(f->pcg_graph); f->pcg_graph = nullptr; }
  if (f->cfg.use_graphs && !f->pcg_graph) {
    hipGraph_t graph = nullptr;
    PM_HIP(hipStreamBeginCapture(f->stream, hipStreamCaptureModeRelaxed));
    for (int k = 0; k < check_every; k++) iterate(k);
    PM_HIP(hipStreamEndCapture(f->stream, &graph));
    const hipError_t e = hipGraphInstantiate(&f->pcg_graph, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) { f->pcg_graph = nullptr; f->err = std::string("graph instantiation failed: ") + hipGetErrorString(e); return 1; }
    f->pcg_graph_x = d_x;
  }
  int it = 0;
  double rr = bb;
  while (it < f->cfg.pcg_max_iter) {
    if (f->cfg.use_graphs && it % check_every == 0 && it + check_every <= f->cfg.pcg_max_iter) {
      PM_HIP(hipGraphLaunch(f->pcg_graph, f->stream));
      it += check_every;
    } else {
      iterate(it);
      it++;
    }
    if (it % check_every == 0 || it == f->cfg.pcg_max_iter) {
      PM_HIP(hipMemcpyAsync(&rr, rr_s, sizeof(double), hipMemcpyDeviceToHost, f->stream));
      PM_HIP(hipStreamSynchronize(f->stream));
      if (!std::isfinite(rr)) { f->err = "Found NaN or Inf in mdu"; *iters = it; return 1; }
      if (rr <= tol2) break;
    }
  }
  *iters = it;
  f->stats.n_pcg_iterations += it;
  if (rr > tol2) { f->err = "PCG did not converge"; return 2; }
  return 0;
}

// PCG on the free DOFs over the stored (upper-block) K with caller-given work memory: r, z, p, q in work
// (4 ndof), partials in red (3 x 2048), scalars in scal (4).  Fixed grids, eager launches (no graph), a
// convergence check every 16 iterations and at the limit; no device memory is allocated.
int pcg_solve_sym(pm_fem* f, const double* d_b, double* d_x, double* work, double* red, double* scal, int* iters) {
  const int64_t np = f->npoint, n = np * f->bs;
  double* r = work;
  double* z = r + n;
  double* p = z + n;
  double* q = p + n;
  const bool lowp = f->cfg.tangent_precision != 0 && f->d_kval32;
  const double rtol2 = f->cfg.pcg_rtol * f->cfg.pcg_rtol;
  const int rb = (int)std::min<int64_t>(2048, std::max<int64_t>(1, (np + PCG_TPB - 1) / PCG_TPB));
  const int rq = (int)std::min<int64_t>(2048, std::max<int64_t>(1, (n + PCG_TPB - 1) / PCG_TPB));
  double* pa = red;         // p.q partials (rq)
  double* pb = red + 2048;  // r.z partials (rb)
  double* pc = red + 4096;  // r.r partials (rb)
  double* rzs = scal;       // rz slots [0], [1]; rr [2]; bb [3]
  double* rr_s = scal + 2;
  double* bb_s = scal + 3;
  *iters = 0;
  if (f->bs == 2) hipLaunchKernelGGL((k_pcg_init<2>), dim3(rb), dim3(PCG_TPB), 0, f->stream, np, d_b, f->d_free, f->d_binv, d_x, r, z, p, pb, pc);
  else hipLaunchKernelGGL((k_pcg_init<3>), dim3(rb), dim3(PCG_TPB), 0, f->stream, np, d_b, f->d_free, f->d_binv, d_x, r, z, p, pb, pc);
  hipLaunchKernelGGL(k_pcg_init_fin, dim3(1), dim3(PCG_TPB), 0, f->stream, rb, pb, pc, rzs, bb_s);
  double bb = 0.0;
  PM_HIP(hipMemcpyAsync(&bb, bb_s, sizeof(double), hipMemcpyDeviceToHost, f->stream));
  PM_HIP(hipStreamSynchronize(f->stream));
  if (bb == 0.0) return 0;
  if (!std::isfinite(bb)) { f->err = "Found NaN or Inf"; return 1; }
  const double tol2 = rtol2 * bb;
  int it = 0;
  double rr = bb;
  while (it < f->cfg.pcg_max_iter) {
    const double* rz_old = rzs + (it & 1);
    double* rz_new = rzs + ((it + 1) & 1);
#define PM_QSYM(BS) do { if (lowp) hipLaunchKernelGGL((k_pcg_q_sym<BS, float>), dim3(rq), dim3(PCG_TPB), 0, f->stream, np, ksym(f, f->d_kval32), f->d_free, p, q, pa); \
                         else hipLaunchKernelGGL((k_pcg_q_sym<BS, double>), dim3(rq), dim3(PCG_TPB), 0, f->stream, np, ksym(f, f->d_kval), f->d_free, p, q, pa); } while (0)
    if (f->bs == 2) {
      PM_QSYM(2);
      hipLaunchKernelGGL((k_pcg_xrz<2>), dim3(rb), dim3(PCG_TPB), 0, f->stream, np, rq, pa, rz_old, p, q, f->d_free, f->d_binv, d_x, r, z, pb, pc);
    } else {
      PM_QSYM(3);
      hipLaunchKernelGGL((k_pcg_xrz<3>), dim3(rb), dim3(PCG_TPB), 0, f->stream, np, rq, pa, rz_old, p, q, f->d_free, f->d_binv, d_x, r, z, pb, pc);
    }
#undef PM_QSYM
    hipLaunchKernelGGL(k_pcg_p2, dim3(rb), dim3(PCG_TPB), 0, f->stream, n, rb, pb, pc, rz_old, rz_new, rr_s, z, p);
    it++;
    if (it % 16 == 0 || it == f->cfg.pcg_max_iter) {
      PM_HIP(hipMemcpyAsync(&rr, rr_s, sizeof(double), hipMemcpyDeviceToHost, f->stream));
      PM_HIP(hipStreamSynchronize(f->stream));
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
void dev_residual(pm_fem* f, double* g, double ell) {
  hipLaunchKernelGGL(k_residual, dim3(nblk(f->ndof)), dim3(TPB), 0, f->stream, f->ndof, f->d_fint, f->d_fext, ell, f->d_rr);
  if (f->nprescribed)
    hipLaunchKernelGGL(k_prescribed_rows, dim3(nblk(f->nprescribed)), dim3(TPB), 0, f->stream, f->nprescribed,
                       f->cfg.lagrange_mult_method, f->d_peq, f->d_lam, f->d_u, f->d_cval, f->d_rr, g);
}

void dev_residual_plain(pm_fem* f, double ell) {
  hipLaunchKernelGGL(k_residual, dim3(nblk(f->ndof)), dim3(TPB), 0, f->stream, f->ndof, f->d_fint, f->d_fext, ell, f->d_rr);
}

// b = (rr - K t) * free with t = g on prescribed DOFs
void dev_rhs(pm_fem* f, const double* g, double* t, double* kt, double* b) {
  dev_fill(f, t, 0.0, f->ndof);
  if (f->nprescribed) hipLaunchKernelGGL(k_scatter_g, dim3(nblk(f->nprescribed)), dim3(TPB), 0, f->stream, f->nprescribed, f->d_peq, g, t);
  dev_spmv(f, t, kt, false);
  hipLaunchKernelGGL(k_rhs, dim3(nblk(f->ndof)), dim3(TPB), 0, f->stream, f->ndof, f->d_rr, kt, f->d_free, b);
}

// x[eq_p] = g_p (g = nullptr: x_p stays as solved, zero) ; y_p = rhs[eq_p] - (K x)[eq_p], rhs = R by default
void dev_complete_correction(pm_fem* f, const double* g, double* x, double* kx, double* y, const double* rhs) {
  if (!f->nprescribed) return;
  if (g) hipLaunchKernelGGL(k_scatter_g, dim3(nblk(f->nprescribed)), dim3(TPB), 0, f->stream, f->nprescribed, f->d_peq, g, x);
  dev_spmv(f, x, kx, false);
  hipLaunchKernelGGL(k_mult_corr, dim3(nblk(f->nprescribed)), dim3(TPB), 0, f->stream, f->nprescribed, f->d_peq,
                     rhs ? rhs : f->d_rr, kx, y);
}

void dev_arc_trial(pm_fem* f, int64_t n, bool extrapolate, double alpha, const double* u_old, const double* u_anc, double* u,
                   double* du) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_arc_trial, dim3(nblk(n)), dim3(TPB), 0, f->stream, n, extrapolate ? 1 : 0, alpha, u_old, u_anc, u, du);
}

void dev_axpy(pm_fem* f, int64_t n, double a, const double* x, double* y) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_axpy, dim3(nblk(n)), dim3(TPB), 0, f->stream, n, a, x, y);
}

void dev_masked(pm_fem* f, const double* x, double* out) {
  hipLaunchKernelGGL(k_masked, dim3(nblk(f->ndof)), dim3(TPB), 0, f->stream, f->ndof, x, f->d_free, out);
}

void dev_update_primary(pm_fem* f, const double* x, const double* y) {
  hipLaunchKernelGGL(k_update_u, dim3(nblk(f->ndof)), dim3(TPB), 0, f->stream, f->ndof, x, f->d_u, f->d_ddu);
  if (f->nprescribed && f->cfg.lagrange_mult_method)
    hipLaunchKernelGGL(k_axpy, dim3(nblk(f->nprescribed)), dim3(TPB), 0, f->stream, f->nprescribed, -1.0, y, f->d_lam);
}

}  // namespace pm
