// Element tangent on the matrix cores (MFMA), FP64 / FP32 / BF16-with-FP32-accumulation.
//
// For D = alpha*I + beta*I(x)I + gamma*s(x)s (Mandel; von_mises.rs:124-167) the element matrix is
//   K_ab = M_ab^T + tr(M_ab) 1 + C_ab,
//   M = sum_q (c alpha/2)_q  g_q g_q^T                (g_q[(a,i)] = G_q[a][i], the gradient vector)
//   C = sum_q (c beta)_q tI_q tI_q^T + (c gamma)_q ts_q ts_q^T (+ (c alpha)_q h_q h_q^T, axisymmetric)
// with tI = B^T I, ts = B^T s and h[(a,0)] = N_a / r.  Both are Gram products over rows built per IP,
// i.e. batched small GEMMs (NN*ND x NN*ND, k = IPs x rows) that run as 16x16 MFMA tiles; only the
// upper tiles are computed.  One wave owns one element; IPs are processed in chunks whose rows are
// staged in LDS (column-major so each lane's operand slice is contiguous).  The epilogue reads the
// upper tiles of M back from LDS, forms K and writes both triangles in rank order (nodes sorted by
// global id), matching the layout consumed by the deterministic scatter.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "pm_internal.hpp"

namespace pm {

namespace {

constexpr double ISQ2 = 0.70710678118654752440;
enum { ERR_SIGD = 1, ERR_NORMS = 2, ERR_DETJ = 4 };

struct MatM {
  int model, plane_stress;
  double K, G, H, z0, E, nu;
};

typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));
typedef short s4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ short to_bf16(float x) {  // round to nearest even
  unsigned u = __float_as_uint(x);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (short)(u >> 16);
}

template <int PREC> struct MfmaT;
template <> struct MfmaT<0> {  // FP64 16x16x4
  typedef double T; typedef d4_t Acc; typedef double Out;
  static constexpr int KS = 4;
  // output register r of lane l holds D[4r + l/16][l%16]
  __device__ static int row(int l, int r) { return 4 * r + (l >> 4); }
};
template <> struct MfmaT<1> {  // FP32 16x16x4
  typedef float T; typedef f4_t Acc; typedef float Out;
  static constexpr int KS = 4;
  __device__ static int row(int l, int r) { return 4 * (l >> 4) + r; }
};
template <> struct MfmaT<2> {  // BF16 16x16x16, FP32 accumulation
  typedef float T; typedef f4_t Acc; typedef float Out;
  static constexpr int KS = 16;
  __device__ static int row(int l, int r) { return 4 * (l >> 4) + r; }
};

template <int ND, int NN, int PREC>
struct MCfg {
  static constexpr int NC = NN * ND;
  static constexpr int NTI = (NC + 15) / 16;
  static constexpr int NCP = NTI * 16;
  static constexpr int NTILE = NTI * (NTI + 1) / 2;
  static constexpr int KS = MfmaT<PREC>::KS;
  static constexpr int IPC = KS;             // IPs per chunk
  static constexpr int QW = 3 + ND * ND + 6 + 1;
};

// LDS layout (per wave = per workgroup): [X (NN*ND) | per-IP data (nip*QW) | union(rows, M tiles)]
template <int ND, int NN, int PREC>
__host__ __device__ constexpr int mfma_rows_stride() {
  // rows per chunk: IPC (M) + 3*IPC (C: tI, ts, h); column-major with an odd stride
  return 4 * MCfg<ND, NN, PREC>::IPC + 1;
}

template <int ND, int NN, int PREC>
__global__ __launch_bounds__(64) void k_tangent_mfma(int64_t nel, int nip, const int32_t* __restrict__ conn,
                                                     const uint8_t* __restrict__ rank, const double* __restrict__ X,
                                                     const double* __restrict__ tab, const double* __restrict__ st,
                                                     typename MfmaT<PREC>::Out* __restrict__ ke, MatM mp, int axisym,
                                                     double thickness, int32_t* err) {
  using C = MCfg<ND, NN, PREC>;
  using MT = MfmaT<PREC>;
  using T = typename MT::T;
  using Acc = typename MT::Acc;
  constexpr int NS = ND == 2 ? 4 : 6;
  constexpr int NC = C::NC, NTI = C::NTI, NCP = C::NCP, NTILE = C::NTILE, KS = C::KS, IPC = C::IPC, QW = C::QW;
  constexpr int RS = mfma_rows_stride<ND, NN, PREC>();  // row-slot stride per column
  constexpr int MS = 17;                                 // M tile row stride
  extern __shared__ double shm[];
  double* sX = shm;
  double* sQ = sX + NN * ND;
  double* uni = sQ + nip * QW;
  T* sL = reinterpret_cast<T*>(uni);          // rows: sL[col * RS + k]
  T* sW = sL + NCP * RS;                      // row weights
  typename MT::Out* sM = reinterpret_cast<typename MT::Out*>(uni);  // epilogue: NTILE upper tiles of M, 16 x MS each
  const int l = threadIdx.x;
  const int64_t e = blockIdx.x;
  if (e >= nel) return;
  const double* __restrict__ tw = tab;
  const double* __restrict__ tN = tab + nip;
  const double* __restrict__ tdN = tN + nip * NN;
  for (int i = l; i < NN; i += 64) {
    const int64_t p = conn[e * NN + i];
#pragma unroll
    for (int d = 0; d < ND; d++) sX[i * ND + d] = X[p * ND + d];
  }
  __syncthreads();
  // J entries of every IP
  for (int w = l; w < nip * ND * ND; w += 64) {
    const int q = w / (ND * ND), ij = w - q * ND * ND, i = ij / ND, j = ij - i * ND;
    const double* dN = tdN + q * NN * ND;
    double s = 0.0;
#pragma unroll 4
    for (int m = 0; m < NN; m++) s += sX[m * ND + i] * dN[m * ND + j];
    sQ[q * QW + 3 + ij] = s;
  }
  __syncthreads();
  const int64_t nipg = nel * nip;
  for (int q = l; q < nip; q += 64) {
    double* Q = sQ + q * QW;
    double J[ND][ND], Ji[ND][ND], det;
#pragma unroll
    for (int i = 0; i < ND; i++)
#pragma unroll
      for (int j = 0; j < ND; j++) J[i][j] = Q[3 + i * ND + j];
    if constexpr (ND == 2) {
      det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
      const double id = 1.0 / det;
      Ji[0][0] = J[1][1] * id; Ji[0][1] = -J[0][1] * id; Ji[1][0] = -J[1][0] * id; Ji[1][1] = J[0][0] * id;
    } else {
      const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
      const double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
      const double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
      det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
      const double id = 1.0 / det;
      Ji[0][0] = c00 * id; Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id; Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
      Ji[1][0] = c01 * id; Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id; Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
      Ji[2][0] = c02 * id; Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id; Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
    }
    if (!(det > 0.0)) *err = ERR_DETJ;
    double r = 0.0;
    if (axisym) for (int m = 0; m < NN; m++) r += tN[q * NN + m] * sX[m * ND];
    const double coef = det * tw[q] * thickness * (axisym ? r : 1.0);
    const int64_t ip = e * nip + q;
    double alpha, beta, gamma = 0.0;
    double s[6] = {0, 0, 0, 0, 0, 0};
    if (mp.plane_stress) {
      alpha = mp.E / (1.0 + mp.nu);
      beta = mp.E * mp.nu / (1.0 - mp.nu * mp.nu);
    } else if (mp.model == 0 || st[(NS + 2) * nipg + ip] != 0.0) {
      alpha = 2.0 * mp.G;
      beta = mp.K - alpha / 3.0;
    } else {
      double sig[NS];
#pragma unroll
      for (int c = 0; c < NS; c++) sig[c] = st[c * nipg + ip];
      const double lam = st[(NS + 1) * nipg + ip];
      const double sm = (sig[0] + sig[1] + sig[2]) / 3.0;
      double ss = 0.0;
#pragma unroll
      for (int c = 0; c < NS; c++) { s[c] = sig[c] - (c < 3 ? sm : 0.0); ss += s[c] * s[c]; }
      const double sd = sqrt(1.5 * ss);
      const double sdtr = sd + lam * 3.0 * mp.G;
      const double norm_s = sd * 0.81649658092772603273;
      if (norm_s < 1e-10) *err = ERR_NORMS;
      const double d = 3.0 * mp.G + mp.H;
      alpha = 2.0 * mp.G * (1.0 - lam * 3.0 * mp.G / sdtr);
      gamma = 6.0 * mp.G * mp.G * (lam / sdtr - 1.0 / d) / (norm_s * norm_s);
      beta = mp.K - alpha / 3.0;
    }
    Q[0] = coef * alpha;
    Q[1] = coef * beta;
    Q[2] = coef * gamma;
#pragma unroll
    for (int i = 0; i < ND; i++)
#pragma unroll
      for (int j = 0; j < ND; j++) Q[3 + i * ND + j] = Ji[i][j];
#pragma unroll
    for (int c = 0; c < 6; c++) Q[3 + ND * ND + c] = s[c];
    Q[3 + ND * ND + 6] = r;
  }
  __syncthreads();
  Acc accM[NTILE], accC[NTILE];
#pragma unroll
  for (int t = 0; t < NTILE; t++) { accM[t] = Acc{0, 0, 0, 0}; accC[t] = Acc{0, 0, 0, 0}; }
  const int nchunk = (nip + IPC - 1) / IPC;
  const int ncr = axisym ? 3 : 2;  // C row kinds: tI, ts (, h)
  for (int ch = 0; ch < nchunk; ch++) {
    // stage the rows of this chunk: k = ql (M), IPC + t*IPC + ql (C, t = 0 tI, 1 ts, 2 h)
    for (int w = l; w < IPC * (NCP / ND + 1); w += 64) {
      const int ql = w / (NCP / ND + 1), m = w - ql * (NCP / ND + 1);
      if (m * ND >= NCP) continue;
      const int q = ch * IPC + ql;
      if (q < nip && m < NN) {
        const double* Q = sQ + q * QW;
        const double* dN = tdN + (q * NN + m) * ND;
        double G[ND];
#pragma unroll
        for (int j = 0; j < ND; j++) {
          double sum = 0.0;
#pragma unroll
          for (int k = 0; k < ND; k++) sum += dN[k] * Q[3 + k * ND + j];
          G[j] = sum;
        }
        const double* sv = Q + 3 + ND * ND;
        const double h = axisym ? tN[q * NN + m] / Q[3 + ND * ND + 6] : 0.0;
        double ts[ND];
        if constexpr (ND == 2) {
          const double sxy = sv[3] * ISQ2;
          ts[0] = sv[0] * G[0] + sxy * G[1] + h * sv[2];
          ts[1] = sxy * G[0] + sv[1] * G[1];
        } else {
          const double sxy = sv[3] * ISQ2, syz = sv[4] * ISQ2, sxz = sv[5] * ISQ2;
          ts[0] = sv[0] * G[0] + sxy * G[1] + sxz * G[2];
          ts[1] = sxy * G[0] + sv[1] * G[1] + syz * G[2];
          ts[2] = sxz * G[0] + syz * G[1] + sv[2] * G[2];
        }
        // rows: g (M), tI = G + h e_x, ts, h e_x
#pragma unroll
        for (int i = 0; i < ND; i++) {
          const int col = m * ND + i;
          sL[col * RS + ql] = (T)G[i];
          sL[col * RS + IPC + ql] = (T)(G[i] + (i == 0 ? h : 0.0));
          sL[col * RS + 2 * IPC + ql] = (T)ts[i];
          sL[col * RS + 3 * IPC + ql] = (i == 0) ? (T)h : T(0);
        }
      } else {
#pragma unroll
        for (int i = 0; i < ND; i++) {
          const int col = m * ND + i;
          if (col >= NCP) break;
          sL[col * RS + ql] = T(0);
          sL[col * RS + IPC + ql] = T(0);
          sL[col * RS + 2 * IPC + ql] = T(0);
          sL[col * RS + 3 * IPC + ql] = T(0);
        }
      }
    }
    for (int w = l; w < IPC; w += 64) {
      const int q = ch * IPC + w;
      const bool ok = q < nip;
      const double* Q = sQ + (ok ? q : 0) * QW;
      sW[w] = ok ? (T)(0.5 * Q[0]) : T(0);     // M rows: c alpha / 2
      sW[IPC + w] = ok ? (T)Q[1] : T(0);       // tI rows: c beta
      sW[2 * IPC + w] = ok ? (T)Q[2] : T(0);   // ts rows: c gamma
      sW[3 * IPC + w] = ok ? (T)Q[0] : T(0);   // h rows: c alpha
    }
    __syncthreads();
    // MFMA: M over rows [0, IPC), C over rows [IPC, IPC*(1+ncr))
    const int col_in = l & 15;
    for (int g = 0; g < 1 + ncr; g++) {
      const int k0 = g * IPC;
      if constexpr (PREC == 2) {
        const int kq = k0 + 4 * (l >> 4);
        s4_t av[NTI], bv[NTI];
#pragma unroll
        for (int I = 0; I < NTI; I++) {
          const T* src = sL + (16 * I + col_in) * RS + kq;
#pragma unroll
          for (int j = 0; j < 4; j++) {
            av[I][j] = to_bf16(src[j]);
            bv[I][j] = to_bf16(src[j] * sW[kq + j]);
          }
        }
        int t = 0;
#pragma unroll
        for (int I = 0; I < NTI; I++)
#pragma unroll
          for (int J = I; J < NTI; J++, t++) {
            if (g == 0) accM[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av[I], bv[J], accM[t], 0, 0, 0);
            else accC[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av[I], bv[J], accC[t], 0, 0, 0);
          }
      } else {
        // KS = 4: the group spans IPC/4 k-steps
#pragma unroll
        for (int ks = 0; ks < IPC / 4; ks++) {
          const int kk = k0 + ks * 4 + (l >> 4);
          T av[NTI], bv[NTI];
          const T wk = sW[kk];
#pragma unroll
          for (int I = 0; I < NTI; I++) { av[I] = sL[(16 * I + col_in) * RS + kk]; bv[I] = av[I] * wk; }
          int t = 0;
#pragma unroll
          for (int I = 0; I < NTI; I++)
#pragma unroll
            for (int J = I; J < NTI; J++, t++) {
              if constexpr (PREC == 0) {
                if (g == 0) accM[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[I], bv[J], accM[t], 0, 0, 0);
                else accC[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[I], bv[J], accC[t], 0, 0, 0);
              } else {
                if (g == 0) accM[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[I], bv[J], accM[t], 0, 0, 0);
                else accC[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[I], bv[J], accC[t], 0, 0, 0);
              }
            }
        }
      }
    }
    __syncthreads();  // rows consumed before the next chunk overwrites them
  }
  // epilogue: M upper tiles to LDS
  {
    int t = 0;
#pragma unroll
    for (int I = 0; I < NTI; I++)
#pragma unroll
      for (int J = I; J < NTI; J++, t++)
#pragma unroll
        for (int r = 0; r < 4; r++) sM[(t * 16 + MT::row(l, r)) * MS + (l & 15)] = accM[t][r];
  }
  __syncthreads();
  auto tile_of = [&](int I, int J) { return I * NTI - I * (I - 1) / 2 + (J - I); };
  // M is symmetric: entries below the tile diagonal are read from their mirror
  auto Mread = [&](int R, int Cc) -> double {
    if ((R >> 4) > (Cc >> 4)) { const int x = R; R = Cc; Cc = x; }
    return (double)sM[(tile_of(R >> 4, Cc >> 4) * 16 + (R & 15)) * MS + (Cc & 15)];
  };
  const uint8_t* rk = rank + e * NN;
  typename MT::Out* out = ke + e * (int64_t)NN * NN * ND * ND;
  int t = 0;
#pragma unroll
  for (int I = 0; I < NTI; I++)
#pragma unroll
    for (int J = I; J < NTI; J++, t++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int R = 16 * I + MT::row(l, r), Cc = 16 * J + (l & 15);
        if (R >= NC || Cc >= NC) continue;
        const int a = R / ND, i = R - a * ND, b = Cc / ND, j = Cc - b * ND;
        if (a > b || (a == b && i > j)) continue;
        double v = (double)accC[t][r] + Mread(ND * a + j, ND * b + i);
        if (i == j) {
          double tr = 0.0;
#pragma unroll
          for (int k = 0; k < ND; k++) tr += Mread(ND * a + k, ND * b + k);
          v += tr;
        }
        const int ra = rk[a], rb = rk[b];
        out[(ra * NN + rb) * ND * ND + i * ND + j] = (typename MT::Out)v;
        if (a != b) out[(rb * NN + ra) * ND * ND + j * ND + i] = (typename MT::Out)v;
        else if (i != j) out[(ra * NN + ra) * ND * ND + j * ND + i] = (typename MT::Out)v;
      }
}

}  // namespace

template <int ND, int NN, int PREC>
static size_t mfma_shm(int nip) {
  using C = MCfg<ND, NN, PREC>;
  using T = typename MfmaT<PREC>::T;
  size_t rows = (size_t)C::NCP * mfma_rows_stride<ND, NN, PREC>() * sizeof(T) + 4 * C::IPC * sizeof(T);
  size_t mt = (size_t)C::NTILE * 16 * 17 * sizeof(typename MfmaT<PREC>::Out);
  return sizeof(double) * ((size_t)NN * ND + (size_t)nip * C::QW) + std::max(rows, mt);
}

// Launches the MFMA tangent for one batch; returns false when the kind has no MFMA variant.
bool launch_tangent_mfma(pm_fem* f, Group& g, int prec) {
  if (g.nel == 0) return true;
  MatM mp{g.param.model, g.plane_stress ? 1 : 0, g.K, g.G, g.param.hh, g.param.z_ini, g.param.young, g.param.poisson};
  const int ax = f->axisym ? 1 : 0;
  const size_t b2 = (size_t)g.ndim * g.ndim;
#define PM_MF(ND, NN)                                                                                                        \
  do {                                                                                                                       \
    if (prec == 0) {                                                                                                         \
      size_t shm = mfma_shm<ND, NN, 0>(g.nip);                                                                               \
      hipLaunchKernelGGL((k_tangent_mfma<ND, NN, 0>), dim3((unsigned)g.nel), dim3(64), shm, f->stream, g.nel, g.nip, g.d_conn, \
                         g.d_rank, f->d_X, g.d_tab, g.d_state, f->d_ke + g.ke_base * b2, mp, ax, f->thickness, f->d_err);    \
    } else {                                                                                                                 \
      if (prec == 1) {                                                                                                       \
        size_t shm = mfma_shm<ND, NN, 1>(g.nip);                                                                             \
        hipLaunchKernelGGL((k_tangent_mfma<ND, NN, 1>), dim3((unsigned)g.nel), dim3(64), shm, f->stream, g.nel, g.nip,       \
                           g.d_conn, g.d_rank, f->d_X, g.d_tab, g.d_state, f->d_ke32 + g.ke_base * b2, mp, ax, f->thickness,  \
                           f->d_err);                                                                                        \
      } else {                                                                                                               \
        size_t shm = mfma_shm<ND, NN, 2>(g.nip);                                                                             \
        hipLaunchKernelGGL((k_tangent_mfma<ND, NN, 2>), dim3((unsigned)g.nel), dim3(64), shm, f->stream, g.nel, g.nip,       \
                           g.d_conn, g.d_rank, f->d_X, g.d_tab, g.d_state, f->d_ke32 + g.ke_base * b2, mp, ax, f->thickness,  \
                           f->d_err);                                                                                        \
      }                                                                                                                      \
    }                                                                                                                        \
  } while (0)
  switch (g.kind) {
    case TRI3: PM_MF(2, 3); break;
    case TRI6: PM_MF(2, 6); break;
    case TRI10: PM_MF(2, 10); break;
    case TRI15: PM_MF(2, 15); break;
    case QUA4: PM_MF(2, 4); break;
    case QUA8: PM_MF(2, 8); break;
    case QUA9: PM_MF(2, 9); break;
    case TET4: PM_MF(3, 4); break;
    case TET10: PM_MF(3, 10); break;
    case HEX8: PM_MF(3, 8); break;
    case HEX20: PM_MF(3, 20); break;
    default: return false;
  }
#undef PM_MF
  return true;
}

}  // namespace pm
