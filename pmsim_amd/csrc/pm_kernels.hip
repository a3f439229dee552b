// HIP kernels (gfx950) of the implicit solid element loop and its linear algebra.
//
// Element kernels work on per-kind batches (template <NDIM, NNODE>); integration-point state is
// structure-of-arrays so consecutive threads touch consecutive addresses.  Assembly into the global
// residual and node-block CSR tangent is a deterministic gather (segmented reduction over a
// precomputed map, contributions in ascending cell order); no atomics are used anywhere.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "pm_internal.hpp"

namespace pm {

namespace {

constexpr double SQ2 = 1.41421356237309504880;
constexpr double ISQ2 = 0.70710678118654752440;
constexpr int TPB = 256;

enum { ERR_SIGD = 1, ERR_NORMS = 2, ERR_DETJ = 4 };

struct Mat {
  int model;         // 0 linear elastic, 1 von Mises
  int plane_stress;  // linear elastic plane stress
  double K, G, H, z0;
  double E, nu;
};

__device__ __forceinline__ void inv_jac2(const double J[2][2], double Ji[2][2], double& det) {
  det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
  double id = 1.0 / det;
  Ji[0][0] = J[1][1] * id; Ji[0][1] = -J[0][1] * id; Ji[1][0] = -J[1][0] * id; Ji[1][1] = J[0][0] * id;
}

__device__ __forceinline__ void inv_jac3(const double J[3][3], double Ji[3][3], double& det) {
  double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
  double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
  double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
  det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
  double id = 1.0 / det;
  Ji[0][0] = c00 * id; Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id; Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
  Ji[1][0] = c01 * id; Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id; Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
  Ji[2][0] = c02 * id; Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id; Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
}

// Jacobian J[i][j] = sum_m X_m[i] dN_m[j], its inverse and determinant.
template <int ND, int NN>
__device__ __forceinline__ void jacobian(const double* __restrict__ Xe, const double* __restrict__ dN, double Ji[ND][ND], double& det) {
  double J[ND][ND];
#pragma unroll
  for (int i = 0; i < ND; i++)
#pragma unroll
    for (int j = 0; j < ND; j++) J[i][j] = 0.0;
#pragma unroll 4
  for (int m = 0; m < NN; m++)
#pragma unroll
    for (int i = 0; i < ND; i++)
#pragma unroll
      for (int j = 0; j < ND; j++) J[i][j] += Xe[m * ND + i] * dN[m * ND + j];
  if constexpr (ND == 2) inv_jac2(J, Ji, det); else inv_jac3(J, Ji, det);
}

// von Mises radial return / linear elastic update on a Mandel stress (src/material/von_mises.rs:170-226,
// linear_elastic.rs:49-71).  Starts from the given state, writes the updated one.
template <int NS>
__device__ __forceinline__ void stress_update(const Mat& mp, double* sig, double& z, double& lam, double& elastic,
                                              double& epbar, const double* de, int32_t* err) {
  if (mp.plane_stress) {
    double c = mp.E / (1.0 - mp.nu * mp.nu);
    sig[0] += c * (de[0] + mp.nu * de[1]);
    sig[1] += c * (mp.nu * de[0] + de[1]);
    sig[3] += c * (1.0 - mp.nu) * de[3];
    return;
  }
  double tr = de[0] + de[1] + de[2];
  double lm = mp.K - 2.0 * mp.G / 3.0;
#pragma unroll
  for (int i = 0; i < NS; i++) sig[i] += 2.0 * mp.G * de[i] + (i < 3 ? lm * tr : 0.0);
  if (mp.model == 0) return;
  elastic = 1.0;
  lam = 0.0;
  double sm = (sig[0] + sig[1] + sig[2]) / 3.0;
  double ss = 0.0;
#pragma unroll
  for (int i = 0; i < NS; i++) { double d = sig[i] - (i < 3 ? sm : 0.0); ss += d * d; }
  double sd = sqrt(1.5 * ss);
  double ftr = sd - z;
  if (ftr / mp.z0 <= 1e-6) return;
  double l = ftr / (3.0 * mp.G + mp.H);
  double beta = 1.0 - l * 3.0 * mp.G / sd;
  double ss2 = 0.0;
#pragma unroll
  for (int i = 0; i < NS; i++) {
    double dev = sig[i] - (i < 3 ? sm : 0.0);
    sig[i] = (i < 3 ? sm : 0.0) + beta * dev;
    ss2 += (beta * dev) * (beta * dev);
  }
  if (sqrt(1.5 * ss2) < 1e-10) *err = ERR_SIGD;
  z += mp.H * l;
  elastic = 0.0;
  lam = l;
  epbar += l;
}

// ------------------------------------------------------------------------------------------------
// update kernel: gather (coords, ddu) -> strain increment -> stress update from the backup state ->
// element internal force f_e = sum_q c_q B_q^T sigma_q   (element_solid.rs:124-138, 190-224)
// ------------------------------------------------------------------------------------------------
template <int ND, int NN>
__global__ __launch_bounds__(TPB) void k_update(int64_t nel, int nip, int epb, const int32_t* __restrict__ conn,
                                                const double* __restrict__ X, const double* __restrict__ ddu,
                                                const double* __restrict__ tab, const double* __restrict__ st_in,
                                                double* __restrict__ st_out, double* __restrict__ fe, Mat mp, int axisym,
                                                double thickness, int do_update, int32_t* err) {
  constexpr int NS = ND == 2 ? 4 : 6;
  constexpr int SW = NS + ND * ND + 1;  // per IP: c*sigma, J^-1, radius
  extern __shared__ double sh[];
  const int ntab = nip * (1 + NN + NN * ND);
  double* sw = sh;
  double* sN = sw + nip;
  double* sdN = sN + nip * NN;
  double* sX = sh + ntab;
  double* sU = sX + epb * NN * ND;
  double* sS = sU + epb * NN * ND;
  const int tid = threadIdx.x;
  const int64_t e0 = (int64_t)blockIdx.x * epb;
  for (int i = tid; i < ntab; i += blockDim.x) sh[i] = tab[i];
  for (int i = tid; i < epb * NN; i += blockDim.x) {
    int64_t e = e0 + i / NN;
    if (e < nel) {
      int64_t p = conn[e0 * NN + i];
#pragma unroll
      for (int d = 0; d < ND; d++) { sX[i * ND + d] = X[p * ND + d]; sU[i * ND + d] = ddu[p * ND + d]; }
    }
  }
  __syncthreads();
  const int64_t nipg = nel * nip;
  for (int w = tid; w < epb * nip; w += blockDim.x) {
    const int el = w / nip, q = w - el * nip;
    const int64_t e = e0 + el;
    if (e >= nel) continue;
    const double* dN = sdN + q * NN * ND;
    const double* Xe = sX + el * NN * ND;
    const double* Ue = sU + el * NN * ND;
    double Ji[ND][ND], det;
    jacobian<ND, NN>(Xe, dN, Ji, det);
    if (!(det > 0.0)) *err = ERR_DETJ;
    double r = 0.0;
    if (axisym) for (int m = 0; m < NN; m++) r += sN[q * NN + m] * Xe[m * ND];
    // strain increment (calculate_strain.rs:18-60), tensor components then Mandel
    double exx = 0, eyy = 0, ezz = 0, exy = 0, eyz = 0, exz = 0;
#pragma unroll 4
    for (int m = 0; m < NN; m++) {
      double G[ND];
#pragma unroll
      for (int j = 0; j < ND; j++) {
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < ND; k++) s += dN[m * ND + k] * Ji[k][j];
        G[j] = s;
      }
      if constexpr (ND == 2) {
        double ux = Ue[m * 2], uy = Ue[m * 2 + 1];
        exx += ux * G[0]; eyy += uy * G[1]; exy += (ux * G[1] + uy * G[0]) * 0.5;
      } else {
        double ux = Ue[m * 3], uy = Ue[m * 3 + 1], uz = Ue[m * 3 + 2];
        exx += ux * G[0]; eyy += uy * G[1]; ezz += uz * G[2];
        exy += (ux * G[1] + uy * G[0]) * 0.5; eyz += (uy * G[2] + uz * G[1]) * 0.5; exz += (ux * G[2] + uz * G[0]) * 0.5;
      }
    }
    if (axisym) for (int m = 0; m < NN; m++) ezz += Ue[m * ND] * sN[q * NN + m] / r;
    double de[NS];
    de[0] = exx; de[1] = eyy; de[2] = ezz; de[3] = exy * SQ2;
    if constexpr (NS == 6) { de[4] = eyz * SQ2; de[5] = exz * SQ2; }
    const int64_t ip = e * nip + q;
    double sig[NS];
#pragma unroll
    for (int c = 0; c < NS; c++) sig[c] = st_in[c * nipg + ip];
    double z = st_in[NS * nipg + ip], lam = st_in[(NS + 1) * nipg + ip], el_ = st_in[(NS + 2) * nipg + ip];
    double ep = st_in[(NS + 3) * nipg + ip];
    if (do_update) {
      stress_update<NS>(mp, sig, z, lam, el_, ep, de, err);
#pragma unroll
      for (int c = 0; c < NS; c++) st_out[c * nipg + ip] = sig[c];
      st_out[NS * nipg + ip] = z;
      st_out[(NS + 1) * nipg + ip] = lam;
      st_out[(NS + 2) * nipg + ip] = el_;
      st_out[(NS + 3) * nipg + ip] = ep;
    }
    double coef = det * sw[q] * thickness * (axisym ? r : 1.0);
    double* S = sS + (el * nip + q) * SW;
#pragma unroll
    for (int c = 0; c < NS; c++) S[c] = coef * sig[c];
#pragma unroll
    for (int i = 0; i < ND; i++)
#pragma unroll
      for (int j = 0; j < ND; j++) S[NS + i * ND + j] = Ji[i][j];
    S[NS + ND * ND] = r;
  }
  __syncthreads();
  for (int w = tid; w < epb * NN; w += blockDim.x) {
    const int el = w / NN, m = w - el * NN;
    const int64_t e = e0 + el;
    if (e >= nel) continue;
    double fm[ND];
#pragma unroll
    for (int d = 0; d < ND; d++) fm[d] = 0.0;
    for (int q = 0; q < nip; q++) {
      const double* S = sS + (el * nip + q) * SW;
      const double* dN = sdN + (q * NN + m) * ND;
      double G[ND];
#pragma unroll
      for (int j = 0; j < ND; j++) {
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < ND; k++) s += dN[k] * S[NS + k * ND + j];
        G[j] = s;
      }
      if constexpr (ND == 2) {
        double sxy = S[3] * ISQ2;
        double fx = S[0] * G[0] + sxy * G[1];
        if (axisym) fx += S[2] * sN[q * NN + m] / S[NS + 4];
        fm[0] += fx;
        fm[1] += sxy * G[0] + S[1] * G[1];
      } else {
        double sxy = S[3] * ISQ2, syz = S[4] * ISQ2, sxz = S[5] * ISQ2;
        fm[0] += S[0] * G[0] + sxy * G[1] + sxz * G[2];
        fm[1] += sxy * G[0] + S[1] * G[1] + syz * G[2];
        fm[2] += sxz * G[0] + syz * G[1] + S[2] * G[2];
      }
    }
#pragma unroll
    for (int d = 0; d < ND; d++) fe[(e * NN + m) * ND + d] = fm[d];
  }
}

// ------------------------------------------------------------------------------------------------
// tangent kernel: K_e node blocks (a <= b) of  sum_q c_q B_a^T D_q B_b  (element_solid.rs:168-185).
// For the isotropic forms used here D = alpha*I + beta*I(x)I + gamma*s(x)s (Mandel), so
//   K_ab = sum_q alpha/2 (Ga.Gb) 1 + alpha/2 Gb Ga^T + alpha h_a h_b e_x e_x^T + beta tIa tIb^T + gamma tsa tsb^T
// with tI = B^T I, ts = B^T s, h = N/r (axisymmetric hoop term).
// ------------------------------------------------------------------------------------------------
template <int ND>
struct TanIP {  // per IP data derived from the state
  double alpha, beta, gamma;
};

template <int ND, int NN, typename KT>
__global__ __launch_bounds__(TPB) void k_tangent(int64_t nel, int nip, int epb, const int32_t* __restrict__ conn,
                                                 const double* __restrict__ X, const double* __restrict__ tab,
                                                 const double* __restrict__ st, KT* __restrict__ ke, Mat mp, int axisym,
                                                 double thickness, int32_t* err) {
  constexpr int NS = ND == 2 ? 4 : 6;
  constexpr int NBLK = NN * (NN + 1) / 2;
  constexpr int PW = 2 * ND + 1;         // per (IP, node): G, ts, h
  constexpr int QW = 3 + ND * ND + 6 + 1;  // per IP: alpha, beta, gamma, J^-1, s (up to 6), r
  extern __shared__ double sh[];
  const int ntab = nip * (1 + NN + NN * ND);
  double* sw = sh;
  double* sN = sw + nip;
  double* sdN = sN + nip * NN;
  double* sX = sh + ntab;
  double* sQ = sX + epb * NN * ND;
  double* sP = sQ + epb * nip * QW;
  __shared__ unsigned char pa[NBLK], pb[NBLK];
  const int tid = threadIdx.x;
  const int64_t e0 = (int64_t)blockIdx.x * epb;
  for (int i = tid; i < ntab; i += blockDim.x) sh[i] = tab[i];
  for (int i = tid; i < NBLK; i += blockDim.x) {
    int a = 0, rem = i;
    while (rem >= NN - a) { rem -= NN - a; a++; }
    pa[i] = (unsigned char)a; pb[i] = (unsigned char)(a + rem);
  }
  for (int i = tid; i < epb * NN; i += blockDim.x) {
    int64_t e = e0 + i / NN;
    if (e < nel) {
      int64_t p = conn[e0 * NN + i];
#pragma unroll
      for (int d = 0; d < ND; d++) sX[i * ND + d] = X[p * ND + d];
    }
  }
  __syncthreads();
  const int64_t nipg = nel * nip;
  // per IP: J^-1, coefficients of D (von_mises.rs:124-167)
  for (int w = tid; w < epb * nip; w += blockDim.x) {
    const int el = w / nip, q = w - el * nip;
    const int64_t e = e0 + el;
    if (e >= nel) continue;
    const double* Xe = sX + el * NN * ND;
    double Ji[ND][ND], det;
    jacobian<ND, NN>(Xe, sdN + q * NN * ND, Ji, det);
    if (!(det > 0.0)) *err = ERR_DETJ;
    double r = 0.0;
    if (axisym) for (int m = 0; m < NN; m++) r += sN[q * NN + m] * Xe[m * ND];
    double coef = det * sw[q] * thickness * (axisym ? r : 1.0);
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
      double lam = st[(NS + 1) * nipg + ip];
      double sm = (sig[0] + sig[1] + sig[2]) / 3.0, ss = 0.0;
#pragma unroll
      for (int c = 0; c < NS; c++) { s[c] = sig[c] - (c < 3 ? sm : 0.0); ss += s[c] * s[c]; }
      double sd = sqrt(1.5 * ss);
      double sdtr = sd + lam * 3.0 * mp.G;
      double norm_s = sd * 0.81649658092772603273;  // sqrt(2/3)
      if (norm_s < 1e-10) *err = ERR_NORMS;
      double d = 3.0 * mp.G + mp.H;
      alpha = 2.0 * mp.G * (1.0 - lam * 3.0 * mp.G / sdtr);
      gamma = 6.0 * mp.G * mp.G * (lam / sdtr - 1.0 / d) / (norm_s * norm_s);
      beta = mp.K - alpha / 3.0;
    }
    double* Q = sQ + (el * nip + q) * QW;
    Q[0] = coef * alpha * 0.5;  // alpha/2 (c folded in)
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
  // per (IP, node): G = dN J^-1, ts = B^T s, h = N / r
  for (int w = tid; w < epb * nip * NN; w += blockDim.x) {
    const int el = w / (nip * NN), rem = w - el * nip * NN, q = rem / NN, m = rem - q * NN;
    if (e0 + el >= nel) continue;
    const double* Q = sQ + (el * nip + q) * QW;
    const double* dN = sdN + (q * NN + m) * ND;
    double G[ND];
#pragma unroll
    for (int j = 0; j < ND; j++) {
      double s = 0.0;
#pragma unroll
      for (int k = 0; k < ND; k++) s += dN[k] * Q[3 + k * ND + j];
      G[j] = s;
    }
    const double* s = Q + 3 + ND * ND;
    double h = axisym ? sN[q * NN + m] / Q[3 + ND * ND + 6] : 0.0;
    double* P = sP + ((el * nip + q) * NN + m) * PW;
    if constexpr (ND == 2) {
      double sxy = s[3] * ISQ2;
      P[0] = G[0]; P[1] = G[1];
      P[2] = s[0] * G[0] + sxy * G[1] + h * s[2];
      P[3] = sxy * G[0] + s[1] * G[1];
    } else {
      double sxy = s[3] * ISQ2, syz = s[4] * ISQ2, sxz = s[5] * ISQ2;
      P[0] = G[0]; P[1] = G[1]; P[2] = G[2];
      P[3] = s[0] * G[0] + sxy * G[1] + sxz * G[2];
      P[4] = sxy * G[0] + s[1] * G[1] + syz * G[2];
      P[5] = sxz * G[0] + syz * G[1] + s[2] * G[2];
    }
    P[2 * ND] = h;
  }
  __syncthreads();
  for (int w = tid; w < epb * NBLK; w += blockDim.x) {
    const int el = w / NBLK, bk = w - el * NBLK;
    const int64_t e = e0 + el;
    if (e >= nel) continue;
    const int a = pa[bk], b = pb[bk];
    double Kb[ND][ND];
#pragma unroll
    for (int i = 0; i < ND; i++)
#pragma unroll
      for (int j = 0; j < ND; j++) Kb[i][j] = 0.0;
    for (int q = 0; q < nip; q++) {
      const double* Q = sQ + (el * nip + q) * QW;
      const double* Pa = sP + ((el * nip + q) * NN + a) * PW;
      const double* Pb = sP + ((el * nip + q) * NN + b) * PW;
      const double ha = Q[0], be = Q[1], ga = Q[2];
      double Ga[ND], Gb[ND], ta[ND], tb[ND];
#pragma unroll
      for (int i = 0; i < ND; i++) { Ga[i] = Pa[i]; Gb[i] = Pb[i]; ta[i] = Pa[ND + i]; tb[i] = Pb[ND + i]; }
      const double hha = Pa[2 * ND], hhb = Pb[2 * ND];
      double dot = 0.0;
#pragma unroll
      for (int i = 0; i < ND; i++) dot += Ga[i] * Gb[i];
      double tIa0 = Ga[0] + hha, tIb0 = Gb[0] + hhb;
#pragma unroll
      for (int i = 0; i < ND; i++) {
        double tIai = i == 0 ? tIa0 : Ga[i];
#pragma unroll
        for (int j = 0; j < ND; j++) {
          double tIbj = j == 0 ? tIb0 : Gb[j];
          double v = ha * Gb[i] * Ga[j] + be * tIai * tIbj + ga * ta[i] * tb[j];
          if (i == j) v += ha * dot;
          Kb[i][j] += v;
        }
      }
      Kb[0][0] += 2.0 * ha * hha * hhb;
    }
    KT* out = ke + ((int64_t)e * NBLK + bk) * ND * ND;
#pragma unroll
    for (int i = 0; i < ND; i++)
#pragma unroll
      for (int j = 0; j < ND; j++) out[i * ND + j] = (KT)Kb[i][j];
  }
}

// deterministic gather of element blocks into node-block CSR values (assembly.rs:124-135)
template <int BS, typename KT>
__global__ __launch_bounds__(TPB) void k_scatter(int64_t nnzb, const int32_t* __restrict__ kptr, const int32_t* __restrict__ kmap,
                                                 const KT* __restrict__ ke, double* __restrict__ kval, float* __restrict__ kval32) {
  constexpr int B2 = BS * BS;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nnzb * B2) return;
  const int64_t blk = t / B2;
  const int c = (int)(t - blk * B2);
  const int ct = (c % BS) * BS + c / BS;
  double s = 0.0;
  for (int32_t k = kptr[blk]; k < kptr[blk + 1]; k++) {
    const int32_t ent = kmap[k];
    const int64_t b = (int64_t)(ent >> 1);
    s += (double)ke[b * B2 + ((ent & 1) ? ct : c)];
  }
  if (kval) kval[t] = s;
  if (kval32) kval32[t] = (float)s;
}

// deterministic gather of element forces into the global internal force vector (assembly.rs:38-46)
template <int ND>
__global__ __launch_bounds__(TPB) void k_rgather(int64_t np, const int32_t* __restrict__ fptr, const int32_t* __restrict__ fmap,
                                                 const double* __restrict__ fe, double* __restrict__ fint) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= np * ND) return;
  const int64_t a = t / ND;
  const int d = (int)(t - a * ND);
  double s = 0.0;
  for (int32_t k = fptr[a]; k < fptr[a + 1]; k++) s += fe[(int64_t)fmap[k] * ND + d];
  fint[t] = s;
}

__global__ void k_init_state(int64_t nipg, int ns, double* st, double z0, int vm) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nipg) return;
  for (int c = 0; c < ns; c++) st[c * nipg + i] = 0.0;
  st[ns * nipg + i] = vm ? z0 : 0.0;
  st[(ns + 1) * nipg + i] = 0.0;
  st[(ns + 2) * nipg + i] = 1.0;
  st[(ns + 3) * nipg + i] = 0.0;
}

__global__ void k_fill(double* x, double v, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = v;
}

__global__ void k_copy(double* __restrict__ d, const double* __restrict__ s, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) d[i] = s[i];
}

}  // namespace

// ================================================================================================
// host-side launchers
// ================================================================================================
static Mat mat_of(const pm_fem* f, const Group& g) {
  Mat m;
  m.model = g.param.model;
  m.plane_stress = g.plane_stress ? 1 : 0;
  m.K = g.K; m.G = g.G; m.H = g.param.hh; m.z0 = g.param.z_ini; m.E = g.param.young; m.nu = g.param.poisson;
  (void)f;
  return m;
}

constexpr size_t SHM_CAP = 60 * 1024;  // dynamic LDS budget per workgroup (static arrays use the rest of 64 KB)

static size_t shm_update(int nd, int nn, int nip, int epb) {
  int ns = nd == 2 ? 4 : 6;
  return sizeof(double) * ((size_t)nip * (1 + nn + nn * nd) + 2 * (size_t)epb * nn * nd + (size_t)epb * nip * (ns + nd * nd + 1));
}
static int epb_update(int nd, int nn, int nip) {
  int epb = std::max(1, TPB / nip);
  while (epb > 1 && shm_update(nd, nn, nip, epb) > SHM_CAP) epb--;
  return epb;
}
static size_t shm_tangent(int nd, int nn, int nip, int epb) {
  return sizeof(double) * ((size_t)nip * (1 + nn + nn * nd) + (size_t)epb * nn * nd + (size_t)epb * nip * (3 + nd * nd + 7) +
                           (size_t)epb * nip * nn * (2 * nd + 1));
}
static int epb_tangent(int nd, int nn, int nip) {
  int epb = std::max(1, TPB / (nn * (nn + 1) / 2));
  while (epb > 1 && shm_tangent(nd, nn, nip, epb) > SHM_CAP) epb--;
  return epb;
}

#define PM_DISPATCH_KIND(KIND, ND, NN, ...)          \
  switch (KIND) {                                    \
    case TRI3: { constexpr int ND = 2, NN = 3; __VA_ARGS__; } break;   \
    case TRI6: { constexpr int ND = 2, NN = 6; __VA_ARGS__; } break;   \
    case TRI10: { constexpr int ND = 2, NN = 10; __VA_ARGS__; } break; \
    case TRI15: { constexpr int ND = 2, NN = 15; __VA_ARGS__; } break; \
    case QUA4: { constexpr int ND = 2, NN = 4; __VA_ARGS__; } break;   \
    case QUA8: { constexpr int ND = 2, NN = 8; __VA_ARGS__; } break;   \
    case QUA9: { constexpr int ND = 2, NN = 9; __VA_ARGS__; } break;   \
    case TET4: { constexpr int ND = 3, NN = 4; __VA_ARGS__; } break;   \
    case TET10: { constexpr int ND = 3, NN = 10; __VA_ARGS__; } break; \
    case HEX8: { constexpr int ND = 3, NN = 8; __VA_ARGS__; } break;   \
    case HEX20: { constexpr int ND = 3, NN = 20; __VA_ARGS__; } break; \
  }

void launch_update(pm_fem* f, Group& g, bool do_update) {
  if (g.nel == 0) return;
  const int epb = epb_update(g.ndim, g.nnode, g.nip);
  const int64_t grid = (g.nel + epb - 1) / epb;
  const size_t shm = shm_update(g.ndim, g.nnode, g.nip, epb);
  const Mat mp = mat_of(f, g);
  PM_DISPATCH_KIND(g.kind, ND, NN,
                   hipLaunchKernelGGL((k_update<ND, NN>), dim3((unsigned)grid), dim3(TPB), shm, f->stream, g.nel, g.nip, epb,
                                      g.d_conn, f->d_X, f->d_ddu, g.d_tab, do_update ? g.d_backup : g.d_state, g.d_state,
                                      f->d_fe + g.fe_base, mp, f->axisym ? 1 : 0, f->thickness, do_update ? 1 : 0, f->d_err));
}

void launch_tangent(pm_fem* f, Group& g) {
  if (g.nel == 0) return;
  const int epb = epb_tangent(g.ndim, g.nnode, g.nip);
  const int64_t grid = (g.nel + epb - 1) / epb;
  const size_t shm = shm_tangent(g.ndim, g.nnode, g.nip, epb);
  const Mat mp = mat_of(f, g);
  const size_t b2 = (size_t)g.ndim * g.ndim;
  if (f->cfg.tangent_precision == 0) {
    PM_DISPATCH_KIND(g.kind, ND, NN,
                     hipLaunchKernelGGL((k_tangent<ND, NN, double>), dim3((unsigned)grid), dim3(TPB), shm, f->stream, g.nel, g.nip,
                                        epb, g.d_conn, f->d_X, g.d_tab, g.d_state, f->d_ke + g.ke_base * b2, mp,
                                        f->axisym ? 1 : 0, f->thickness, f->d_err));
  } else {
    PM_DISPATCH_KIND(g.kind, ND, NN,
                     hipLaunchKernelGGL((k_tangent<ND, NN, float>), dim3((unsigned)grid), dim3(TPB), shm, f->stream, g.nel, g.nip,
                                        epb, g.d_conn, f->d_X, g.d_tab, g.d_state, f->d_ke32 + g.ke_base * b2, mp,
                                        f->axisym ? 1 : 0, f->thickness, f->d_err));
  }
}

void launch_scatter(pm_fem* f) {
  const int64_t n = f->nnzb * f->bs * f->bs;
  const unsigned grid = (unsigned)((n + TPB - 1) / TPB);
  const bool lowp = f->cfg.tangent_precision != 0;
  if (f->bs == 2) {
    if (!lowp) hipLaunchKernelGGL((k_scatter<2, double>), dim3(grid), dim3(TPB), 0, f->stream, f->nnzb, f->d_kmap_ptr, f->d_kmap, f->d_ke, f->d_kval, (float*)nullptr);
    else hipLaunchKernelGGL((k_scatter<2, float>), dim3(grid), dim3(TPB), 0, f->stream, f->nnzb, f->d_kmap_ptr, f->d_kmap, f->d_ke32, f->d_kval, f->d_kval32);
  } else {
    if (!lowp) hipLaunchKernelGGL((k_scatter<3, double>), dim3(grid), dim3(TPB), 0, f->stream, f->nnzb, f->d_kmap_ptr, f->d_kmap, f->d_ke, f->d_kval, (float*)nullptr);
    else hipLaunchKernelGGL((k_scatter<3, float>), dim3(grid), dim3(TPB), 0, f->stream, f->nnzb, f->d_kmap_ptr, f->d_kmap, f->d_ke32, f->d_kval, f->d_kval32);
  }
}

void launch_residual_gather(pm_fem* f) {
  const int64_t n = f->npoint * f->ndim;
  const unsigned grid = (unsigned)((n + TPB - 1) / TPB);
  if (f->ndim == 2) hipLaunchKernelGGL((k_rgather<2>), dim3(grid), dim3(TPB), 0, f->stream, f->npoint, f->d_fmap_ptr, f->d_fmap, f->d_fe, f->d_fint);
  else hipLaunchKernelGGL((k_rgather<3>), dim3(grid), dim3(TPB), 0, f->stream, f->npoint, f->d_fmap_ptr, f->d_fmap, f->d_fe, f->d_fint);
}

void launch_init_state(pm_fem* f, Group& g) {
  const int64_t n = g.nel * g.nip;
  if (n == 0) return;
  hipLaunchKernelGGL(k_init_state, dim3((unsigned)((n + TPB - 1) / TPB)), dim3(TPB), 0, f->stream, n, f->nsig, g.d_state,
                     g.param.z_ini, g.param.model == 1 ? 1 : 0);
}

void launch_reset_lambda(pm_fem* f, Group& g) {
  const int64_t n = g.nel * g.nip;
  if (n == 0) return;
  hipLaunchKernelGGL(k_fill, dim3((unsigned)((n + TPB - 1) / TPB)), dim3(TPB), 0, f->stream, g.d_state + (int64_t)(f->nsig + 1) * n,
                     0.0, n);
}

void launch_copy(pm_fem* f, double* dst, const double* src, int64_t n) {
  if (n == 0) return;
  int64_t blocks = std::min<int64_t>((n + TPB - 1) / TPB, 8192);
  hipLaunchKernelGGL(k_copy, dim3((unsigned)blocks), dim3(TPB), 0, f->stream, dst, src, n);
}

void dev_fill(pm_fem* f, double* x, double v, int64_t n) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_fill, dim3((unsigned)((n + TPB - 1) / TPB)), dim3(TPB), 0, f->stream, x, v, n);
}

int check_device_error(pm_fem* f) {
  int32_t e = 0;
  PM_HIP(hipMemcpyAsync(&e, f->d_err, sizeof(int32_t), hipMemcpyDeviceToHost, f->stream));
  PM_HIP(hipStreamSynchronize(f->stream));
  if (e == 0) return 0;
  int32_t z = 0;
  PM_HIP(hipMemcpyAsync(f->d_err, &z, sizeof(int32_t), hipMemcpyHostToDevice, f->stream));
  PM_HIP(hipStreamSynchronize(f->stream));
  if (e & ERR_DETJ) f->err = "negative or zero Jacobian determinant";
  else if (e & ERR_SIGD) f->err = "von Mises plastic update must not lead to zero sigma_d (sigma_d < 1e-10)";
  else f->err = "von Mises stiffness cannot be computed with zero deviatoric norm (norm_s < 1e-10)";
  return 1;
}

size_t update_shared_bytes(const Group& g) { return shm_update(g.ndim, g.nnode, g.nip, epb_update(g.ndim, g.nnode, g.nip)); }
size_t tangent_shared_bytes(const Group& g) { return shm_tangent(g.ndim, g.nnode, g.nip, epb_tangent(g.ndim, g.nnode, g.nip)); }

}  // namespace pm
