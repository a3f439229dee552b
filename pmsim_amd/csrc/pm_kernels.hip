// HIP kernels (gfx950) of the implicit solid element loop and its linear algebra.
//
// Element kernels work on per-kind batches (template <NDIM, NNODE>); integration-point state is
// structure-of-arrays so consecutive threads touch consecutive addresses.  Assembly into the global
// residual and node-block CSR tangent is a deterministic gather (segmented reduction over a
// precomputed map, contributions in ascending cell order); no atomics are used anywhere.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <tuple>
#include <type_traits>
#include <vector>

#include "pm_internal.hpp"

namespace pm {

namespace {

constexpr double SQ2 = 1.41421356237309504880;
constexpr double ISQ2 = 0.70710678118654752440;
constexpr int TPB = 256;

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
                                              double& epbar, const double* de, int32_t* err, uint64_t key) {
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
  if (sqrt(1.5 * ss2) < 1e-10) raise_error(err, ERR_SIGD, key);
  z += mp.H * l;
  elastic = 0.0;
  lam = l;
  epbar += l;
}

// Workgroups are dispatched round-robin over the 8 XCDs (each with its own L2); remap so each XCD
// walks a contiguous range of logical blocks and neighbouring elements / rows share one L2.
__device__ __forceinline__ int64_t xcd_block() {
  const int64_t nb = gridDim.x, b = blockIdx.x;
  const int64_t per = nb >> 3, rem = nb & 7, x = b & 7, i = b >> 3;
  return x * per + (x < rem ? x : rem) + i;
}

// kernel arguments of one element group (or one handle, for the gathers)
struct UpdArgs {
  int64_t nel; int nip, epb;
  const int32_t* conn; const double* geo; const double* ddu; const double* tab; const double* st_in;
  double* st_out; double* fe; Mat mp; int axisym; double thickness; int do_update; int32_t* err; uint64_t ekey;
};
template <typename KT>
struct TanArgs {
  int64_t nel; int nip, ch;
  const uint8_t* rank; const double* geo; const double* tab; const double* st; KT* ke;
  Mat mp; int axisym; double thickness; int32_t* err; uint64_t ekey;
  int cap;  // output staging per pass, in KT values (a multiple of 4)
};
template <typename KT>
struct ScatArgs { int64_t nnzb; const int32_t* korder; const int32_t* kptr; const int32_t* kmap; const KT* ke; double* kval; float* kval32; };
template <typename KT>
struct ScatWArgs { int64_t nwave; const int32_t* kwave; const int32_t* kptr; const int32_t* kmap; const KT* ke; double* kval; float* kval32; };
struct RgArgs { int64_t np; const int32_t* fptr; const int32_t* fmap; const double* fe; double* fint; };

typedef __attribute__((address_space(3))) void lds_void_k;

// the value held by the neighbouring lane of this lane's pair (lanes 2k and 2k + 1; DPP quad permutation
// [1,0,3,2] on both words of the double)
__device__ __forceinline__ double pair_lane(double v) {
  const int lo = __double2loint(v), hi = __double2hiint(v);
  return __hiloint2double(__builtin_amdgcn_mov_dpp(hi, 0xB1, 0xf, 0xf, false), __builtin_amdgcn_mov_dpp(lo, 0xB1, 0xf, 0xf, false));
}

// ------------------------------------------------------------------------------------------------
// update kernel: gather of the displacement increments -> strain increment -> stress update from the
// backup state -> element internal force f_e = sum_q c_q B_q^T sigma_q   (element_solid.rs:124-138, 190-224)
// One IP per thread, epb elements per workgroup.  The workgroup's IP geometry (contiguous in HBM) and the
// shape-function derivatives are copied into LDS by direct global->LDS loads, all issued before the first
// wait (the derivatives row by row, with an odd stride per IP, so that the lanes of one element, which read
// different IPs, hit different banks).  The displacement gradient is H J^-1 with H = sum_m u_m dN_m^T.  For
// the force loop each IP leaves T = c sigma J^-T (and the hoop term c sigma_zz / r, axisymmetric) in place
// of its geometry, so f_e[m] = sum_q T_q dN_mq: one thread per element node, ND*ND FMA per IP.
// ------------------------------------------------------------------------------------------------
// f_e[m] = sum_q T_q dN_mq over an element block's IPs (T = c sigma J^-T, and the hoop term, in the IP's
// geometry slot).  PAIR (rules of eight or more IPs): each thread takes two nodes of an element (m and
// m + NH, so every T read serves both) over one half of the IPs; the halves sit in neighbouring lanes and
// are added once, lower half first.  Otherwise one thread per node over all IPs.
template <int ND, int NN, bool PAIR>
__device__ __forceinline__ void force_loop(const double* sdN, const double* sG, const double* __restrict__ tN,
                                           double* __restrict__ fe, int64_t e0, int nloc, int nip, int axisym) {
  constexpr int GW = ND * ND + 2, DS = NN * ND + 1;
  constexpr int NH = PAIR ? (NN + 1) / 2 : NN;  // node groups per element
  constexpr int SPLIT = PAIR ? 2 : 1;
  const int qh = PAIR ? (nip + 1) / 2 : nip;
  const int nwork = nloc * NH * SPLIT;  // even when PAIR (the lanes of a pair are active together)
  for (int w = threadIdx.x; w < nwork; w += blockDim.x) {
    const int h = PAIR ? (w & 1) : 0;
    const int pw = PAIR ? (w >> 1) : w;
    const int el = pw / NH, m0 = pw - el * NH;
    const bool two = PAIR && ((NN % 2 == 0) || m0 + NH < NN);
    const int m1 = two ? m0 + NH : m0;
    const int64_t e = e0 + el;
    double f0[ND], f1[ND];
#pragma unroll
    for (int d = 0; d < ND; d++) f0[d] = f1[d] = 0.0;
    const int qa = h ? qh : 0, qb = h ? nip : qh;
    const double* T = sG + (el * nip + qa) * GW;
    for (int q = qa; q < qb; q++, T += GW) {
      const double* dN0 = sdN + q * DS + m0 * ND;
      const double* dN1 = sdN + q * DS + m1 * ND;
#pragma unroll
      for (int i = 0; i < ND; i++)
#pragma unroll
        for (int k = 0; k < ND; k++) {
          const double t = T[i * ND + k];
          f0[i] += t * dN0[k];
          if constexpr (PAIR) f1[i] += t * dN1[k];
        }
      if (axisym) {
        f0[0] += T[ND * ND] * tN[q * NN + m0];
        if constexpr (PAIR) f1[0] += T[ND * ND] * tN[q * NN + m1];
      }
    }
    if constexpr (PAIR) {
#pragma unroll
      for (int d = 0; d < ND; d++) {
        const double g0 = pair_lane(f0[d]), g1 = pair_lane(f1[d]);
        f0[d] += g0;  // the lower half's lane: its half plus the upper one
        f1[d] += g1;
      }
    }
    if (h == 0) {
#pragma unroll
      for (int d = 0; d < ND; d++) fe[(e * NN + m0) * ND + d] = f0[d];
      if (two)
#pragma unroll
        for (int d = 0; d < ND; d++) fe[(e * NN + m1) * ND + d] = f1[d];
    }
  }
}

template <int ND, int NN>
__device__ __forceinline__ void update_body(int64_t blk, const UpdArgs& A) {
  const int64_t nel = A.nel;
  const int nip = A.nip, epb = A.epb;
  const int32_t* __restrict__ conn = A.conn;
  const double* __restrict__ geo = A.geo;
  const double* __restrict__ ddu = A.ddu;
  const double* __restrict__ tab = A.tab;
  const double* __restrict__ st_in = A.st_in;
  double* __restrict__ st_out = A.st_out;
  double* __restrict__ fe = A.fe;
  const Mat mp = A.mp;
  const int axisym = A.axisym, do_update = A.do_update;
  const double thickness = A.thickness;
  int32_t* err = A.err;
  const uint64_t ekey = A.ekey;
  constexpr int NS = ND == 2 ? 4 : 6;
  constexpr int GW = ND * ND + 2;      // per IP in geo: J^-1, det J, radius
  constexpr int DS = NN * ND + 1;      // LDS stride of the derivatives of one IP (odd)
  static_assert(ND * ND + 1 <= GW, "T and the hoop term fit the IP's geometry slot");
  extern __shared__ double sh[];
  const double* __restrict__ tw = tab;
  const double* __restrict__ tN = tab + nip;
  const double* __restrict__ tdN = tN + nip * NN;
  double* sdN = sh;                    // [nip][DS]
  double* sU = sdN + nip * DS;         // [epb][NN][ND] displacement increments of the block's nodes
  double* sG = sU + epb * NN * ND;     // [epb * nip][GW] geometry, then T and the hoop term
  const int tid = threadIdx.x;
  const int64_t e0 = blk * epb;
  const int nloc = (int)min((int64_t)epb, nel - e0);
  const int64_t nipg = nel * nip;
  // one IP per thread (epb * nip <= blockDim); its state is requested first, so that the loads are in flight
  // while the workgroup's tables arrive
  const int el = tid / nip, q = tid - el * nip;
  const bool mine = tid < nloc * nip;
  const int64_t ip = mine ? (e0 + el) * nip + q : 0;
  double sig[NS];
#pragma unroll
  for (int c = 0; c < NS; c++) sig[c] = mine ? st_in[c * nipg + ip] : 0.0;
  double z = mine ? st_in[NS * nipg + ip] : 0.0, lam = mine ? st_in[(NS + 1) * nipg + ip] : 0.0;
  double el_ = mine ? st_in[(NS + 2) * nipg + ip] : 0.0, ep = mine ? st_in[(NS + 3) * nipg + ip] : 0.0;
  {
    // geometry: 2 ng dwords in loads of one dword per lane (TPB per round); a wavefront's load lands at its
    // own 64-dword slice of the round
    const uint32_t* __restrict__ gsrc = reinterpret_cast<const uint32_t*>(geo + e0 * nip * GW);
    uint32_t* gdst = reinterpret_cast<uint32_t*>(sG);
    const int nd2 = 2 * nloc * nip * GW, wv = tid >> 6, ln = tid & 63;
    static_assert(TPB % 64 == 0, "whole wavefronts");
    for (int r = 0; r < 2 * GW; r++) {  // 2 ng <= 2 TPB GW dwords (epb nip <= TPB)
      const int base = r * TPB + wv * 64;  // uniform per wavefront
      if (r * TPB >= nd2) break;
      if (base + ln < nd2) __builtin_amdgcn_global_load_lds(gsrc + base + ln, (lds_void_k*)(gdst + base), 4, 0, 0);
    }
    // derivatives: IP q's row of NN ND doubles to sdN + q DS, in slices of 64 dwords
    constexpr int RW = 2 * NN * ND, RS = (RW + 63) / 64;  // dwords per row, slices per row
    const uint32_t* __restrict__ dsrc = reinterpret_cast<const uint32_t*>(tdN);
    uint32_t* ddst = reinterpret_cast<uint32_t*>(sdN);
    for (int t = wv; t < nip * RS; t += TPB / 64) {  // uniform per wavefront
      const int q = t / RS, o = (t - q * RS) * 64;
      if (o + ln < RW) __builtin_amdgcn_global_load_lds(dsrc + q * RW + o + ln, (lds_void_k*)(ddst + 2 * q * DS + o), 4, 0, 0);
    }
    for (int i = tid; i < nloc * NN; i += blockDim.x) {
      const int64_t p = conn[e0 * NN + i];
#pragma unroll
      for (int d = 0; d < ND; d++) sU[i * ND + d] = ddu[p * ND + d];
    }
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");  // the direct loads (not tracked by the compiler)
  }
  __syncthreads();
  if (mine) {
    // J^-1, det J and radius of the IP (formed once for the fixed mesh, k_geometry)
    double* gq = sG + tid * GW;
    double Ji[ND][ND];
#pragma unroll
    for (int i = 0; i < ND; i++)
#pragma unroll
      for (int j = 0; j < ND; j++) Ji[i][j] = gq[i * ND + j];
    const double det = gq[ND * ND], r = gq[ND * ND + 1];
    const double* dN = sdN + q * DS;
    const double* Ue = sU + el * NN * ND;
    if (!(det > 0.0)) raise_error(err, ERR_DETJ, ekey + (uint64_t)ip);
    // strain increment (calculate_strain.rs:18-60): displacement gradient L = H J^-1 with
    // H = sum_m u_m dN_m^T (dN with respect to the reference coordinates), then the tensor components and
    // Mandel; the shear sums 2 e_ij = L_ij + L_ji are halved once
    double H[ND][ND];
#pragma unroll
    for (int i = 0; i < ND; i++)
#pragma unroll
      for (int k = 0; k < ND; k++) H[i][k] = 0.0;
#pragma unroll 4
    for (int m = 0; m < NN; m++)
#pragma unroll
      for (int i = 0; i < ND; i++)
#pragma unroll
        for (int k = 0; k < ND; k++) H[i][k] = fma(Ue[m * ND + i], dN[m * ND + k], H[i][k]);
    double L[ND][ND];
#pragma unroll
    for (int i = 0; i < ND; i++)
#pragma unroll
      for (int j = 0; j < ND; j++) {
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < ND; k++) s = fma(H[i][k], Ji[k][j], s);
        L[i][j] = s;
      }
    double exx = L[0][0], eyy = L[1][1], ezz = 0, exy = L[0][1] + L[1][0], eyz = 0, exz = 0;
    if constexpr (ND == 3) { ezz = L[2][2]; eyz = L[1][2] + L[2][1]; exz = L[0][2] + L[2][0]; }
    if (axisym) for (int m = 0; m < NN; m++) ezz += Ue[m * ND] * tN[q * NN + m] / r;
    double de[NS];
    de[0] = exx; de[1] = eyy; de[2] = ezz; de[3] = (exy * 0.5) * SQ2;
    if constexpr (NS == 6) { de[4] = (eyz * 0.5) * SQ2; de[5] = (exz * 0.5) * SQ2; }
    if (do_update) {
      stress_update<NS>(mp, sig, z, lam, el_, ep, de, err, ekey + (uint64_t)ip);
#pragma unroll
      for (int c = 0; c < NS; c++) st_out[c * nipg + ip] = sig[c];
      st_out[NS * nipg + ip] = z;
      st_out[(NS + 1) * nipg + ip] = lam;
      st_out[(NS + 2) * nipg + ip] = el_;
      st_out[(NS + 3) * nipg + ip] = ep;
    }
    const double coef = det * tw[q] * thickness * (axisym ? r : 1.0);
    // c sigma as a tensor (the Mandel shear components divided by sqrt 2), then T = c sigma J^-T
    double S[ND][ND];
    S[0][0] = coef * sig[0]; S[1][1] = coef * sig[1];
    S[0][1] = S[1][0] = (coef * sig[3]) * ISQ2;
    if constexpr (ND == 3) {
      S[2][2] = coef * sig[2];
      S[1][2] = S[2][1] = (coef * sig[4]) * ISQ2;
      S[0][2] = S[2][0] = (coef * sig[5]) * ISQ2;
    }
#pragma unroll
    for (int i = 0; i < ND; i++)
#pragma unroll
      for (int k = 0; k < ND; k++) {
        double t = 0.0;
#pragma unroll
        for (int j = 0; j < ND; j++) t += S[i][j] * Ji[k][j];
        gq[i * ND + k] = t;
      }
    gq[ND * ND] = axisym ? (coef * sig[2]) / r : 0.0;
  }
  __syncthreads();
  if (nip >= 8) force_loop<ND, NN, true>(sdN, sG, tN, fe, e0, nloc, nip, axisym);  // uniform
  else force_loop<ND, NN, false>(sdN, sG, tN, fe, e0, nloc, nip, axisym);
}

// ------------------------------------------------------------------------------------------------
// tangent kernel: K_e node blocks of  sum_q c_q B_a^T D_q B_b  (element_solid.rs:168-185).
// For the isotropic forms used here D = alpha*I + beta*I(x)I + gamma*s(x)s (Mandel), so
//   K_ab = sum_q alpha/2 (Ga.Gb) 1 + alpha/2 Gb Ga^T + alpha h_a h_b e_x e_x^T + beta tIa tIb^T + gamma tsa tsb^T
// with tI = B^T I, ts = B^T s, h = N/r (axisymmetric hoop term): three rank-1 terms and a diagonal per
// node block (33 FMA per block and IP instead of the dense 6x6 triple product).
//
// Work split: one wavefront per workgroup (TW threads), EPW elements x NT upper 2x2 tiles of node blocks,
// one thread per tile, so no phase waits on another wavefront.  The D coefficients of every IP of the
// workgroup's elements are formed first from the state and the stored J^-1 and det J; the IPs are then
// taken in chunks of ch: the node vectors (G, h, ts) of a chunk are staged in LDS, node m in the slot of
// its rank (nodes sorted by global id), and every thread adds the chunk's contributions to its tile in
// registers, so no reduction across threads is needed.  The element matrices, contiguous in memory, are
// finally staged in LDS (over the dead IP data, in at most a few passes) as their upper node blocks in rank
// order (ke_block_offset) and written with 16-byte stores.
// ------------------------------------------------------------------------------------------------
constexpr int TW = 64;  // threads per tangent workgroup: one wavefront

template <int ND>
struct NodeVec { static constexpr int W = 2 * ND + 1; };  // G(ND), h, ts(ND): odd stride in dwords/2 -> no LDS bank conflicts

template <int ND, int NN>
struct TanShape {
  static constexpr int NH = (NN + 1) / 2;
  static constexpr int NT = NH * (NH + 1) / 2;
  static constexpr int NNP = 2 * NH;
  static constexpr int PW = NodeVec<ND>::W;
  static constexpr int QW = 3 + ND * ND + 6 + 1;  // alpha/2, beta, gamma (times c), J^-1, s, r
  static constexpr int B2 = ND * ND;
  static constexpr int EPW = TW / NT;             // elements per workgroup
  static_assert(NT <= TW, "one element's tiles fit one wavefront");
};

// elements per tangent workgroup of a kind with nn nodes (TanShape::EPW)
inline int tangent_epw(int nn) { const int nh = (nn + 1) / 2; return TW / (nh * (nh + 1) / 2); }

template <int ND, int NN, typename KT>
__device__ __forceinline__ void tangent_body(int64_t blk, const TanArgs<KT>& A) {
  const int64_t nel = A.nel;
  const int nip = A.nip, ch = A.ch, cap = A.cap;
  const uint8_t* __restrict__ rank = A.rank;
  const double* __restrict__ geo = A.geo;
  const double* __restrict__ tab = A.tab;
  const double* __restrict__ st = A.st;
  KT* __restrict__ ke = A.ke;
  const Mat mp = A.mp;
  const int axisym = A.axisym;
  const double thickness = A.thickness;
  int32_t* err = A.err;
  const uint64_t ekey = A.ekey;
  using TS = TanShape<ND, NN>;
  constexpr int NS = ND == 2 ? 4 : 6;
  constexpr int NT = TS::NT, NNP = TS::NNP, PW = TS::PW, QW = TS::QW, B2 = TS::B2, NH = TS::NH, EPW = TS::EPW;
  constexpr int GW = ND * ND + 2;               // per IP in geo: J^-1, det J, radius
  constexpr int EW = packed_ke_values(NN, ND);  // values per stored element matrix (padded)
  constexpr int EU = NN * (NN + 1) / 2 * B2;    // values in use
  constexpr int SR = ke_split_row(NN, ND);
  constexpr int PADS = SR < NN ? ke_block_offset(NN, ND, SR, SR) - 1 : -1;  // padding value between the halves
  extern __shared__ __attribute__((aligned(16))) double sh[];
  double* sQ = sh;                              // EPW * nip * QW: per-IP coefficients
  // node vectors of one IP chunk, ES doubles per element (odd, so that the reads of neighbouring elements
  // fall in disjoint LDS bank pairs)
  const int ES = (ch * NNP * PW) | 1;
  double* sP = sQ + EPW * nip * QW;
  __shared__ uint8_t srank[EPW * NN];
  const int tid = threadIdx.x;
  const int64_t e0 = blk * EPW;
  const double* __restrict__ tN = tab + nip;      // shape functions and derivatives, read through the L1
  const double* __restrict__ tdN = tN + nip * NN;
  for (int i = tid; i < EPW * NN; i += TW)
    if (e0 + i / NN < nel) srank[i] = rank[e0 * NN + i];
  const int64_t nipg = nel * nip;
  // per IP: D coefficients (von_mises.rs:124-167), deviator, J^-1 (formed once for the fixed mesh)
  for (int w = tid; w < EPW * nip; w += TW) {
    const int el = EPW == 1 ? 0 : w / nip, q = w - el * nip;
    const int64_t e = e0 + el;
    if (e >= nel) continue;
    double* Q = sQ + w * QW;
    const int64_t ip = e * nip + q;
    const double* gq = geo + ip * GW;
    const double det = gq[ND * ND], r = gq[ND * ND + 1];
    if (!(det > 0.0)) raise_error(err, ERR_DETJ, ekey + (uint64_t)ip);
    double coef = det * tab[q] * thickness * (axisym ? r : 1.0);
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
      if (norm_s < 1e-10) raise_error(err, ERR_NORMS, ekey + (uint64_t)ip);
      double d = 3.0 * mp.G + mp.H;
      alpha = 2.0 * mp.G * (1.0 - lam * 3.0 * mp.G / sdtr);
      gamma = 6.0 * mp.G * mp.G * (lam / sdtr - 1.0 / d) / (norm_s * norm_s);
      beta = mp.K - alpha / 3.0;
    }
    Q[0] = coef * alpha * 0.5;
    Q[1] = coef * beta;
    Q[2] = coef * gamma;
#pragma unroll
    for (int ij = 0; ij < ND * ND; ij++) Q[3 + ij] = gq[ij];
#pragma unroll
    for (int c = 0; c < 6; c++) Q[3 + ND * ND + c] = s[c];
    Q[3 + ND * ND + 6] = r;
  }
  // this thread: element el, tile (I <= J) -> blocks (2I+u, 2J+v)
  const int el = tid / NT, tile = tid - el * NT;
  const bool active = el < EPW && (e0 + el) < nel;
  int I = 0, J = 0;
  {
    int rem = tile;
    while (rem >= NH - I) { rem -= NH - I; I++; }
    J = I + rem;
  }
  double K[2][2][ND][ND];
  double Dg[2][2];  // sum over the IPs of alpha/2 (Ga . Gb): added to the block diagonals after the loop
#pragma unroll
  for (int u = 0; u < 2; u++)
#pragma unroll
    for (int v = 0; v < 2; v++) {
      Dg[u][v] = 0.0;
#pragma unroll
      for (int i = 0; i < ND; i++)
#pragma unroll
        for (int j = 0; j < ND; j++) K[u][v][i][j] = 0.0;
    }
  // per IP and 2x2 tile: three rank-1 updates per block (27 FMA in 3D) and the scaled dot product; the
  // hoop terms only when axisymmetric (the loop is specialised so plane problems carry none of them)
  auto accumulate = [&](auto axc, int q0, int nq) {
    constexpr bool AX = decltype(axc)::value;
#pragma unroll 1
    for (int ql = 0; ql < nq; ql++) {
      const double* Q = sQ + (el * nip + q0 + ql) * QW;
      const double ha = Q[0], be = Q[1], ga = Q[2];
      const double* Pq = sP + el * ES + ql * NNP * PW;
      double Ga[2][ND], tIa[2][ND], hA[2];
#pragma unroll
      for (int u = 0; u < 2; u++) {
        const double* Pa = Pq + (2 * I + u) * PW;
#pragma unroll
        for (int i = 0; i < ND; i++) { Ga[u][i] = Pa[i]; tIa[u][i] = Ga[u][i]; }
        hA[u] = AX ? Pa[ND] : 0.0;
        if (AX) tIa[u][0] += hA[u];
      }
      // alpha and beta terms (and the axisymmetric hoop term): chained accumulations, one FMA each into the
      // accumulator (a parenthesised sum would cost a multiply, two FMAs and an add)
#pragma unroll
      for (int v = 0; v < 2; v++) {
        const double* Pb = Pq + (2 * J + v) * PW;
        double ub[ND], vb[ND];
#pragma unroll
        for (int i = 0; i < ND; i++) {
          ub[i] = ha * Pb[i];
          vb[i] = be * Pb[i];
        }
        double hh = 0.0;
        if (AX) { const double hb = Pb[ND]; vb[0] = fma(be, hb, vb[0]); hh = 2.0 * ha * hb; }
#pragma unroll
        for (int u = 0; u < 2; u++) {
#pragma unroll
          for (int i = 0; i < ND; i++) {
            Dg[u][v] = fma(Ga[u][i], ub[i], Dg[u][v]);
#pragma unroll
            for (int j = 0; j < ND; j++) {
              double k = K[u][v][i][j];
              k = fma(ub[i], Ga[u][j], k);
              k = fma(tIa[u][i], vb[j], k);
              K[u][v][i][j] = k;
            }
          }
          if (AX) K[u][v][0][0] = fma(hA[u], hh, K[u][v][0][0]);
        }
      }
      // the s(x)s term only where the IP is plastic (gamma = 0 at elastic IPs, so skipping it adds nothing): a
      // branch that only adds to the accumulators (uniform over the lanes of one element and IP); the ts
      // vectors are read from LDS only here
      if (ga != 0.0) {
        double ta[2][ND];
#pragma unroll
        for (int u = 0; u < 2; u++)
#pragma unroll
          for (int i = 0; i < ND; i++) ta[u][i] = Pq[(2 * I + u) * PW + ND + 1 + i];
#pragma unroll
        for (int v = 0; v < 2; v++) {
          const double* Pb = Pq + (2 * J + v) * PW;
          double wb[ND];
#pragma unroll
          for (int i = 0; i < ND; i++) wb[i] = ga * Pb[ND + 1 + i];
#pragma unroll
          for (int u = 0; u < 2; u++)
#pragma unroll
            for (int i = 0; i < ND; i++)
#pragma unroll
              for (int j = 0; j < ND; j++) K[u][v][i][j] = fma(ta[u][i], wb[j], K[u][v][i][j]);
        }
      }
    }
  };
  // the chunk loop is specialised as a whole (the branch on the uniform axisym flag sits outside it, so
  // only one variant's loop invariants are live)
  auto chunks = [&](auto axc) {
  for (int q0 = 0; q0 < nip; q0 += ch) {
    const int nq = min(ch, nip - q0);
    __syncthreads();  // coefficients ready (first chunk); the previous chunk's node vectors consumed
    // node vectors of every (IP of the chunk, element, node), node m in the slot of its rank (the tiles
    // and the stored blocks are in rank order); the padding slot of odd NN stays zero.  One item at a time:
    // the tile accumulators are live here
#pragma unroll 1
    for (int w = tid; w < nq * EPW * NNP; w += TW) {
      const int m = w % NNP, r1 = w / NNP, e = r1 % EPW, ql = r1 / EPW;
      const bool pad = e0 + e >= nel || m >= NN;
      double* P = sP + e * ES + (ql * NNP + (pad ? m : srank[e * NN + m])) * PW;
      if (pad) {
#pragma unroll
        for (int k = 0; k < PW; k++) P[k] = 0.0;
        continue;
      }
      const int q = q0 + ql;
      const double* Q = sQ + (e * nip + q) * QW;
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
      double h = axisym ? tN[q * NN + m] / Q[3 + ND * ND + 6] : 0.0;
      if constexpr (ND == 2) {
        double sxy = sv[3] * ISQ2;
        P[0] = G[0]; P[1] = G[1]; P[2] = h;
        P[3] = sv[0] * G[0] + sxy * G[1] + h * sv[2];
        P[4] = sxy * G[0] + sv[1] * G[1];
      } else {
        double sxy = sv[3] * ISQ2, syz = sv[4] * ISQ2, sxz = sv[5] * ISQ2;
        P[0] = G[0]; P[1] = G[1]; P[2] = G[2]; P[3] = h;
        P[4] = sv[0] * G[0] + sxy * G[1] + sxz * G[2];
        P[5] = sxy * G[0] + sv[1] * G[1] + syz * G[2];
        P[6] = sxz * G[0] + syz * G[1] + sv[2] * G[2];
      }
    }
    __syncthreads();
    if (active) accumulate(axc, q0, nq);
  }
  };
  if (axisym) chunks(std::true_type{});
  else chunks(std::false_type{});
#pragma unroll
  for (int u = 0; u < 2; u++)
#pragma unroll
    for (int v = 0; v < 2; v++)
#pragma unroll
      for (int i = 0; i < ND; i++) K[u][v][i][i] += Dg[u][v];
  // the workgroup's element matrices are contiguous: staged in LDS over the IP data, cap values per pass,
  // then copied out with 16-byte stores.  The tiles are in rank order, so a tile's upper blocks are stored
  // blocks as they are.  A diagonal node block takes its lower half from the upper one, so every element
  // matrix (and K) is exactly symmetric (the local check of assembly.rs:78-90 holds by construction)
  KT* stage = reinterpret_cast<KT*>(sh);
  typedef KT V __attribute__((ext_vector_type(16 / sizeof(KT))));
  constexpr int VL = 16 / (int)sizeof(KT);
  static_assert(EW % VL == 0, "element matrix is a whole number of 16-byte pieces");
  const int nval = (int)min((int64_t)EPW, nel - e0) * EW;
  for (int p0 = 0; p0 < nval; p0 += cap) {
    __syncthreads();  // node vectors / the previous pass consumed
    if (active) {
      const int base = el * EW - p0;
#pragma unroll
      for (int u = 0; u < 2; u++)
#pragma unroll
        for (int v = 0; v < 2; v++) {
          const int a = 2 * I + u, b = 2 * J + v;
          if (a >= NN || b >= NN || a > b) continue;
          const int o = base + ke_block_offset(NN, ND, a, b);
#pragma unroll
          for (int i = 0; i < ND; i++)
#pragma unroll
            for (int j = 0; j < ND; j++) {
              const int x = o + i * ND + j;
              if ((unsigned)x < (unsigned)cap) stage[x] = (KT)(a == b && j < i ? K[u][v][j][i] : K[u][v][i][j]);
            }
        }
      if (tile == 0) {  // padding values: the one between the halves (Hex20) and the tail
        if (PADS >= 0 && (unsigned)(base + PADS) < (unsigned)cap) stage[base + PADS] = (KT)0;
        for (int i = EU + (PADS >= 0 ? 1 : 0); i < EW; i++)
          if ((unsigned)(base + i) < (unsigned)cap) stage[base + i] = (KT)0;
      }
    }
    __syncthreads();
    const int nv = min(cap, nval - p0) / VL;
    V* dst = reinterpret_cast<V*>(ke + e0 * EW + p0);
    const V* src = reinterpret_cast<const V*>(stage);
    for (int i = tid; i < nv; i += TW) dst[i] = src[i];
  }
}

// deterministic gather of element blocks into node-block CSR values (assembly.rs:124-135): one thread
// per stored (upper) CSR block sums its contributing element blocks in ascending cell order.  A map
// entry is the value offset of a stored (upper, rank-order) element block; with the sign bit set the
// block would be read transposed (the maps for the upper triangle never set it).
template <int BS, typename KT>
__device__ __forceinline__ void scatter_body(int64_t blk_order, const ScatArgs<KT>& A) {
  const int64_t nnzb = A.nnzb;
  const int32_t* __restrict__ korder = A.korder;
  const int32_t* __restrict__ kptr = A.kptr;
  const int32_t* __restrict__ kmap = A.kmap;
  const KT* __restrict__ ke = A.ke;
  double* __restrict__ kval = A.kval;
  float* __restrict__ kval32 = A.kval32;
  constexpr int B2 = BS * BS;
  constexpr int PF = 2;  // contributions whose map entries and blocks are in flight together
  __shared__ double stage[TPB * B2];
  const int64_t b0 = (int64_t)korder[blk_order] * blockDim.x;  // chunks in their processing order
  const int64_t blk = b0 + threadIdx.x;
  double acc[B2];
#pragma unroll
  for (int c = 0; c < B2; c++) acc[c] = 0.0;
  if (blk < nnzb) {
    const int32_t k1 = kptr[blk + 1];
    for (int32_t k = kptr[blk]; k < k1; k += PF) {
      int32_t idx[PF];
#pragma unroll
      for (int j = 0; j < PF; j++) idx[j] = (k + j < k1) ? kmap[k + j] : 0;
      double v[PF][B2];
#pragma unroll
      for (int j = 0; j < PF; j++) {
        const KT* src = ke + (idx[j] & 0x7fffffff);
        KT raw[B2];
#pragma unroll
        for (int c = 0; c < B2; c++) raw[c] = src[c];  // contiguous: the address does not depend on the flag
        const bool tr = idx[j] < 0;
#pragma unroll
        for (int r = 0; r < BS; r++)
#pragma unroll
          for (int c = 0; c < BS; c++) v[j][r * BS + c] = (double)(tr ? raw[c * BS + r] : raw[r * BS + c]);
      }
      // ascending contribution order
#pragma unroll
      for (int j = 0; j < PF; j++)
        if (k + j < k1) {
#pragma unroll
          for (int c = 0; c < B2; c++) acc[c] += v[j][c];
        }
    }
  }
  // stage the workgroup's blocks and write them out contiguously
#pragma unroll
  for (int c = 0; c < B2; c++) stage[threadIdx.x * B2 + c] = acc[c];
  __syncthreads();
  const int64_t nb = min((int64_t)blockDim.x, nnzb - b0);
  if (nb <= 0) return;
  const int64_t nv = nb * B2;
  if (kval) {
    double* dst = kval + b0 * B2;
    for (int64_t i = threadIdx.x; i < nv; i += blockDim.x) dst[i] = stage[i];
  }
  if (kval32) {
    float* dst = kval32 + b0 * B2;
    for (int64_t i = threadIdx.x; i < nv; i += blockDim.x) dst[i] = (float)stage[i];
  }
}

// Balanced K gather: a wavefront per range of whole CSR blocks (at most 64 contributions and 64 blocks,
// host-packed; a block with more contributions has its own range, taken in rounds of 64).  Every lane loads
// one contribution (map entry and element block) and stages it in the wavefront's LDS; the lane that owns a
// block then sums that block's staged contributions in ascending order from zero (the sums of scatter_body,
// bit for bit), and the range's blocks are written contiguously.  No lane waits for another's contribution
// list, so all of a range's loads are in flight together.  Wavefronts never wait for each other (no
// barriers); LDS accesses of one wavefront complete in order.
template <int BS, typename KT>
__device__ __forceinline__ void scatter_wave_body(int64_t w, const ScatWArgs<KT>& A) {
  constexpr int B2 = BS * BS;
  __shared__ double stage[TPB * B2];
  const int lane = threadIdx.x & 63;
  if (w >= A.nwave) return;  // uniform over the wavefront
  double* st = stage + (threadIdx.x >> 6) * 64 * B2;
  const int32_t* __restrict__ kptr = A.kptr;
  const int32_t* __restrict__ kmap = A.kmap;
  const KT* __restrict__ ke = A.ke;
  const int32_t b0 = A.kwave[w], nb = A.kwave[w + 1] - b0;  // nb <= 64
  const int32_t c0 = kptr[b0], c1 = kptr[b0 + nb];
  int32_t s = 0, e = 0;  // owner lane: its block's contributions
  if (lane < nb) { s = kptr[b0 + lane]; e = kptr[b0 + lane + 1]; }
  double acc[B2];
#pragma unroll
  for (int c = 0; c < B2; c++) acc[c] = 0.0;
  for (int32_t r0 = c0; r0 < c1; r0 += 64) {  // one round unless the range is a single large block
    const int32_t k = r0 + lane;
    if (k < c1) {
      const int32_t idx = kmap[k];
      const KT* src = ke + (idx & 0x7fffffff);
      KT raw[B2];
#pragma unroll
      for (int c = 0; c < B2; c++) raw[c] = src[c];
      const bool tr = idx < 0;
#pragma unroll
      for (int r = 0; r < BS; r++)
#pragma unroll
        for (int c = 0; c < BS; c++) st[lane * B2 + r * BS + c] = (double)(tr ? raw[c * BS + r] : raw[r * BS + c]);
    }
    asm volatile("" ::: "memory");
    if (lane < nb) {
      const int32_t lo = max(s, r0), hi = min(e, r0 + 64);
      for (int32_t t = lo; t < hi; t++)  // ascending contribution order
#pragma unroll
        for (int c = 0; c < B2; c++) acc[c] += st[(t - r0) * B2 + c];
    }
    asm volatile("" ::: "memory");
  }
  if (lane < nb)
#pragma unroll
    for (int c = 0; c < B2; c++) st[lane * B2 + c] = acc[c];
  asm volatile("" ::: "memory");
  const int nv = nb * B2;
  if (A.kval) {
    double* dst = A.kval + (int64_t)b0 * B2;
    for (int i = lane; i < nv; i += 64) dst[i] = st[i];
  }
  if (A.kval32) {
    float* dst = A.kval32 + (int64_t)b0 * B2;
    for (int i = lane; i < nv; i += 64) dst[i] = (float)st[i];
  }
}

// deterministic gather of element forces into the global internal force vector (assembly.rs:38-46)
template <int ND>
__device__ __forceinline__ void rgather_body(int64_t blk, const RgArgs& A) {
  const int32_t* __restrict__ fptr = A.fptr;
  const int32_t* __restrict__ fmap = A.fmap;
  const double* __restrict__ fe = A.fe;
  double* __restrict__ fint = A.fint;
  const int64_t t = blk * blockDim.x + threadIdx.x;
  if (t >= A.np * ND) return;
  const int64_t a = t / ND;
  const int d = (int)(t - a * ND);
  // contributions four at a time: their map entries, then their values, requested together (the sum
  // still runs in ascending order)
  double s = 0.0;
  const int32_t k1 = fptr[a + 1];
  for (int32_t k = fptr[a]; k < k1; k += 4) {
    int32_t m[4];
#pragma unroll
    for (int j = 0; j < 4; j++) m[j] = fmap[min(k + j, k1 - 1)];
    double v[4];
#pragma unroll
    for (int j = 0; j < 4; j++) v[j] = fe[(int64_t)m[j] * ND + d];
#pragma unroll
    for (int j = 0; j < 4; j++)
      if (k + j < k1) s += v[j];
  }
  fint[t] = s;
}

// one launch per kernel for a handle (workgroups remapped per XCD), and batched launches over the element
// groups of several handles: map[blockIdx.x] = job << 20 | workgroup of that job (the per-element and per-row
// arithmetic is the same either way, so are the bits)
__device__ __forceinline__ void unpack_job(int32_t e, int& job, int& w) { job = e >> 20; w = e & ((1 << 20) - 1); }

template <int ND, int NN>
__global__ __launch_bounds__(TPB) void k_update(UpdArgs A) { update_body<ND, NN>(xcd_block(), A); }
template <int ND, int NN>
__global__ __launch_bounds__(TPB) void kb_update(const UpdArgs* __restrict__ jobs, const int32_t* __restrict__ map) {
  int j, w;
  unpack_job(map[blockIdx.x], j, w);
  update_body<ND, NN>(w, jobs[j]);
}
template <int ND, int NN, typename KT>
__global__ __launch_bounds__(TW) void k_tangent(TanArgs<KT> A) { tangent_body<ND, NN, KT>(xcd_block(), A); }
template <int ND, int NN, typename KT>
__global__ __launch_bounds__(TW) void kb_tangent(const TanArgs<KT>* __restrict__ jobs, const int32_t* __restrict__ map) {
  int j, w;
  unpack_job(map[blockIdx.x], j, w);
  tangent_body<ND, NN, KT>(w, jobs[j]);
}
template <int BS, typename KT>
__global__ __launch_bounds__(TPB) void k_scatter(ScatArgs<KT> A) { scatter_body<BS, KT>(xcd_block(), A); }
template <int BS, typename KT>
__global__ __launch_bounds__(TPB) void k_scatter_w(ScatWArgs<KT> A) {
  scatter_wave_body<BS, KT>(xcd_block() * (TPB / 64) + (threadIdx.x >> 6), A);
}
template <int BS, typename KT>
__global__ __launch_bounds__(TPB) void kb_scatter(const ScatArgs<KT>* __restrict__ jobs, const int32_t* __restrict__ map) {
  int j, w;
  unpack_job(map[blockIdx.x], j, w);
  scatter_body<BS, KT>(w, jobs[j]);
}
template <int ND>
__global__ __launch_bounds__(TPB) void k_rgather(RgArgs A) { rgather_body<ND>(xcd_block(), A); }
template <int ND>
__global__ __launch_bounds__(TPB) void kb_rgather(const RgArgs* __restrict__ jobs, const int32_t* __restrict__ map) {
  int j, w;
  unpack_job(map[blockIdx.x], j, w);
  rgather_body<ND>(w, jobs[j]);
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

// J^-1, det J and the radius (axisymmetric) of every IP of a group, one IP per thread: the same arithmetic
// as the update kernel's (jacobian), formed once since the mesh does not move (small strain)
template <int ND, int NN>
__global__ __launch_bounds__(TPB) void k_geometry(int64_t nel, int nip, const int32_t* __restrict__ conn,
                                                  const double* __restrict__ X, const double* __restrict__ tab, int axisym,
                                                  double* __restrict__ geo) {
  const int64_t ip = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (ip >= nel * nip) return;
  const int64_t e = ip / nip;
  const int q = (int)(ip - e * nip);
  double Xe[NN * ND];
#pragma unroll
  for (int m = 0; m < NN; m++) {
    const int64_t p = conn[e * NN + m];
#pragma unroll
    for (int d = 0; d < ND; d++) Xe[m * ND + d] = X[p * ND + d];
  }
  const double* tN = tab + nip;
  double Ji[ND][ND], det;
  jacobian<ND, NN>(Xe, tN + nip * NN + q * NN * ND, Ji, det);
  double r = 0.0;
  if (axisym) for (int m = 0; m < NN; m++) r += tN[q * NN + m] * Xe[m * ND];
  double* o = geo + ip * (ND * ND + 2);
#pragma unroll
  for (int i = 0; i < ND; i++)
#pragma unroll
    for (int j = 0; j < ND; j++) o[i * ND + j] = Ji[i][j];
  o[ND * ND] = det;
  o[ND * ND + 1] = r;
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

// error key base of a group's IPs (raise_error)
static uint64_t group_key(const pm_fem* f, const Group& g) { return (uint64_t)(&g - f->groups.data()) << kErrKeyShift; }

constexpr size_t SHM_CAP = 60 * 1024;  // dynamic LDS budget per workgroup (static arrays use the rest of 64 KB)

static size_t shm_update(int nd, int nn, int nip, int epb) {
  return sizeof(double) * ((size_t)nip * (nn * nd + 1) + (size_t)epb * nn * nd + (size_t)epb * nip * (nd * nd + 2));
}
static int epb_update(int nd, int nn, int nip) {
  int epb = std::max(1, TPB / nip);  // the update kernel takes one IP per thread: epb * nip <= TPB (nip <= 27 here)
  while (epb > 1 && shm_update(nd, nn, nip, epb) > SHM_CAP) epb--;
  return epb;
}
// dynamic LDS of a vector-tangent workgroup (bytes): per-IP coefficients and one chunk's node vectors, and
// at least half of the workgroup's element matrices for the output staging (at most two passes)
static size_t shm_tangent(int nd, int nn, int nip, int ch, int kt_bytes) {
  const int epw = tangent_epw(nn), nh = (nn + 1) / 2, pw = 2 * nd + 1, qw = 3 + nd * nd + 7;
  const size_t base = (size_t)epw * nip * qw + (size_t)epw * ((ch * 2 * nh * pw) | 1);
  const size_t half = ((size_t)epw * packed_ke_values(nn, nd) / 2 + 3) / 4 * 4;  // values
  const size_t dbl = std::max(base, (half * kt_bytes + 7) / 8);
  return sizeof(double) * ((dbl + 1) / 2 * 2);
}
// output staging per pass of a vector-tangent workgroup, in values (a multiple of 4)
static int stage_cap(size_t shm, int kt_bytes) { return (int)(shm / kt_bytes) / 4 * 4; }
constexpr size_t TANGENT_LDS_TARGET = 13 * 1024;  // twelve one-wavefront workgroups per CU (160 KB LDS)
// IPs per chunk of the vector tangent: the fewest rounds of node-vector staging (one round: a pass of the
// wavefront over its items; each chunk also costs a synchronisation) within the LDS target
static int split_tangent(int nd, int nn, int nip, int kt_bytes) {
  const int epw = tangent_epw(nn), nnp = 2 * ((nn + 1) / 2);
  int best = 1, best_cost = 1 << 30;
  for (int ch = 1; ch <= std::min(nip, 16); ch++) {
    if (ch > 1 && shm_tangent(nd, nn, nip, ch, kt_bytes) > TANGENT_LDS_TARGET) break;
    int cost = 0;
    for (int q0 = 0; q0 < nip; q0 += ch) cost += (std::min(ch, nip - q0) * epw * nnp + TW - 1) / TW + 1;
    if (cost <= best_cost) { best_cost = cost; best = ch; }
  }
  return best;
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

static UpdArgs upd_args(pm_fem* f, Group& g, bool do_update, int epb) {
  UpdArgs A;
  A.nel = g.nel; A.nip = g.nip; A.epb = epb;
  A.conn = g.d_conn; A.geo = g.d_geo; A.ddu = f->d_ddu; A.tab = g.d_tab;
  A.st_in = do_update ? g.d_backup : g.d_state; A.st_out = g.d_state; A.fe = f->d_fe + g.fe_base;
  A.mp = mat_of(f, g); A.axisym = f->axisym ? 1 : 0; A.thickness = f->thickness; A.do_update = do_update ? 1 : 0;
  A.err = f->d_err; A.ekey = group_key(f, g);
  return A;
}

// von Mises under general plasticity: pmsim's Elastoplastic model has no consistent stiffness
// (elastoplastic.rs:528-536 returns "TODO"), so no tangent is formed; its ODE stress update is not built here
constexpr const char* kGpStiffness = "TODO";
constexpr const char* kGpUpdate = "the general-plasticity (Elastoplastic) stress update is not available on the device";
static void refuse(pm_fem* f, const char* msg) { f->launch_refused = true; f->refuse_msg = msg; }

// the refusal above, taken before a pass is launched or captured: true (with f->err) when the pass would
// form a tangent, or update the stress, of a general-plasticity group
bool gp_refuse(pm_fem* f, bool do_update, bool tangent) {
  for (const auto& g : f->groups) {
    if (!g.general_plasticity || g.nel == 0) continue;
    if (tangent) { f->err = kGpStiffness; return true; }
    if (do_update) { f->err = kGpUpdate; return true; }
  }
  return false;
}

void launch_update(pm_fem* f, Group& g, bool do_update) {
  if (g.nel == 0) return;
  if (do_update && g.general_plasticity) { refuse(f, kGpUpdate); return; }
  const int epb = epb_update(g.ndim, g.nnode, g.nip);
  const int64_t grid = (g.nel + epb - 1) / epb;
  const size_t shm = shm_update(g.ndim, g.nnode, g.nip, epb);
  const UpdArgs A = upd_args(f, g, do_update, epb);
  PM_DISPATCH_KIND(g.kind, ND, NN, hipLaunchKernelGGL((k_update<ND, NN>), dim3((unsigned)grid), dim3(TPB), shm, f->stream, A));
}

// tangent kernel of a group: the tuned or forced choice, else the default (matrix cores for BF16, the vector
// kernel otherwise).  1: vector kernel (the persistent Hex20 form where it fits, tangent_hex20_fits), 2: matrix
// cores (for FP32 and BF16 the persistent Hex20 form where it fits, tangent_hex20_mfma_fits), 4: vector kernel
// with one workgroup per element group, 5: the per-group matrix-core kernel (never the persistent forms)
static int tangent_kernel_of(const pm_fem* f, const Group& g) {
  const int prec = f->cfg.tangent_precision;
  if (prec == 2) return g.tune_kernel == 5 ? 5 : 2;  // BF16 runs on the matrix cores only
  if (g.tune_kernel) return g.tune_kernel;
  // FP64 and FP32: the vector kernel.  For FP32 its FP64 products rounded once on store keep an inexact
  // Newton solve closest to the FP64 one (the matrix-core form, FP32 accumulation, is faster per pass but
  // no faster per solve, which the PCG dominates; `mfma` selects it)
  return 1;
}
// true when a group's tangent runs as the persistent Hex20 kernel
static bool runs_hex20(const pm_fem* f, const Group& g) {
  return tangent_kernel_of(f, g) == 1 && g.tune_ch == 0 && tangent_hex20_fits(f, g);
}
// a group's matrix-core tangent: the persistent Hex20 kernel where it fits and is not excluded, else the
// per-group kernel; false when neither has a variant for the group
static bool launch_matrix_core(pm_fem* f, Group& g) {
  if (tangent_kernel_of(f, g) == 2 && tangent_hex20_mfma_fits(f, g) && launch_tangent_hex20_mfma(f, g)) return true;
  return launch_tangent_mfma(f, g, f->cfg.tangent_precision);
}

// IPs per chunk of a group's vector tangent: the tuned value, else the heuristic
static int tangent_split(const pm_fem* f, const Group& g) {
  if (g.tune_ch > 0) return g.tune_ch;
  return split_tangent(g.ndim, g.nnode, g.nip, f->cfg.tangent_precision == 0 ? 8 : 4);
}

template <typename KT>
static TanArgs<KT> tan_args(pm_fem* f, Group& g, int ch, KT* ke, size_t shm) {
  TanArgs<KT> A;
  A.nel = g.nel; A.nip = g.nip; A.ch = ch; A.cap = stage_cap(shm, (int)sizeof(KT));
  A.rank = g.d_rank; A.geo = g.d_geo; A.tab = g.d_tab; A.st = g.d_state; A.ke = ke + g.ke_base;
  A.mp = mat_of(f, g); A.axisym = f->axisym ? 1 : 0; A.thickness = f->thickness; A.err = f->d_err; A.ekey = group_key(f, g);
  return A;
}

void launch_tangent(pm_fem* f, Group& g) {
  if (g.nel == 0) return;
  if (g.general_plasticity) { refuse(f, kGpStiffness); return; }
  const int kern = tangent_kernel_of(f, g);
  if ((kern == 2 || kern == 5) && launch_matrix_core(f, g)) return;
  if (runs_hex20(f, g) && launch_tangent_hex20(f, g)) return;
  const int ch = tangent_split(f, g), epw = tangent_epw(g.nnode);
  const int64_t grid = (g.nel + epw - 1) / epw;
  if (!ke_range_ok(g.ke_base, g.nel, g.kew, packed_ke_values(g.nnode, g.ndim), f->n_ke_values)) {
    f->launch_refused = true;  // reported by check_device_error; never launched out of range
    return;
  }
  if (f->cfg.tangent_precision == 0) {
    const size_t shm = shm_tangent(g.ndim, g.nnode, g.nip, ch, 8);
    const TanArgs<double> A = tan_args(f, g, ch, f->d_ke, shm);
    PM_DISPATCH_KIND(g.kind, ND, NN, hipLaunchKernelGGL((k_tangent<ND, NN, double>), dim3((unsigned)grid), dim3(TW), shm, f->stream, A));
  } else {
    const size_t shm = shm_tangent(g.ndim, g.nnode, g.nip, ch, 4);
    const TanArgs<float> A = tan_args(f, g, ch, f->d_ke32, shm);
    PM_DISPATCH_KIND(g.kind, ND, NN, hipLaunchKernelGGL((k_tangent<ND, NN, float>), dim3((unsigned)grid), dim3(TW), shm, f->stream, A));
  }
}

// Tangent kernel choice per group.  Modes 1, 2 and 4 force a kernel; mode 3 times every candidate (the
// matrix-core kernel, the persistent Hex20 vector kernel where it fits and the vector kernel under each
// feasible work split) on the current state and keeps the fastest; mode 0 keeps the default.
int tune_tangent(pm_fem* f) {
  const int mode = f->cfg.tangent_kernel;
  release_graphs(f);
  for (auto& g : f->groups) { g.tune_kernel = (mode == 1 || mode == 2 || mode == 4 || mode == 5) ? mode : 0; g.tune_ch = 0; }
  if (mode != 3) return 0;
  struct Cand { int kern, ch; };
  hipEvent_t ev0, ev1;
  PM_HIP(hipEventCreate(&ev0));
  PM_HIP(hipEventCreate(&ev1));
  int rc = 0;
  for (auto& g : f->groups) {
    if (g.nel == 0 || g.general_plasticity) continue;  // no tangent to time
    std::vector<Cand> cands;
    cands.push_back({2, 0});
    if (f->cfg.tangent_precision != 0 && tangent_hex20_mfma_fits(f, g)) cands.push_back({5, 0});
    if (f->cfg.tangent_precision != 2 && tangent_hex20_fits(f, g)) cands.push_back({1, 0});
    if (f->cfg.tangent_precision != 2)
      for (int ch : {1, 2, 3, 4, 6, 9, 14}) {
        if (ch > g.nip) continue;
        if (shm_tangent(g.ndim, g.nnode, g.nip, ch, f->cfg.tangent_precision == 0 ? 8 : 4) > SHM_CAP) continue;
        cands.push_back({4, ch});
      }
    float best = 1e30f;
    Cand pick = cands[0];
    for (const Cand& c : cands) {
      g.tune_kernel = c.kern; g.tune_ch = c.ch;
      launch_tangent(f, g);  // warm-up
      if (hipEventRecord(ev0, f->stream) != hipSuccess) { rc = 1; break; }
      for (int r = 0; r < 3; r++) launch_tangent(f, g);
      if (hipEventRecord(ev1, f->stream) != hipSuccess || hipEventSynchronize(ev1) != hipSuccess) { rc = 1; break; }
      float ms = 0.0f;
      (void)hipEventElapsedTime(&ms, ev0, ev1);
      if (f->cfg.verbose & 1)
        std::fprintf(stderr, "tangent tuning: kind %d, %s kernel, %d IPs per chunk: %.3f ms\n", g.kind,
                     c.kern == 2 ? "matrix-core" : c.kern == 5 ? "per-group matrix-core" : (c.kern == 1 ? "persistent vector" : "vector"),
                     c.ch, ms / 3.0f);
      if (ms < best) { best = ms; pick = c; }
    }
    g.tune_kernel = pick.kern; g.tune_ch = pick.ch;
    if (rc) break;
  }
  (void)hipEventDestroy(ev0);
  (void)hipEventDestroy(ev1);
  if (rc) { f->err = "tangent tuning failed: " + std::string(hipGetErrorString(hipGetLastError())); return 1; }
  return check_device_error(f);
}

static_assert(TPB == GATHER_CHUNK, "the gather's chunk order assumes one workgroup of TPB blocks per chunk");
// the gather forms the stored (upper) node blocks of K only
template <typename KT>
static ScatArgs<KT> scat_args(const pm_fem* f, const KT* ke) {
  ScatArgs<KT> A;
  A.nnzb = f->nnzu; A.korder = f->d_korder; A.kptr = f->d_kmap_ptr; A.kmap = f->d_kmap; A.ke = ke;
  A.kval = f->cfg.tangent_precision == 0 ? f->d_kval : nullptr;
  A.kval32 = f->cfg.tangent_precision == 0 ? nullptr : f->d_kval32;
  return A;
}

static RgArgs rg_args(const pm_fem* f) {
  RgArgs A;
  A.np = f->npoint; A.fptr = f->d_fmap_ptr; A.fmap = f->d_fmap; A.fe = f->d_fe; A.fint = f->d_fint;
  return A;
}

template <typename KT>
static ScatWArgs<KT> scatw_args(const pm_fem* f, const KT* ke) {
  ScatWArgs<KT> A;
  A.nwave = f->nkwave; A.kwave = f->d_kwave; A.kptr = f->d_kmap_ptr; A.kmap = f->d_kmap; A.ke = ke;
  A.kval = f->cfg.tangent_precision == 0 ? f->d_kval : nullptr;
  A.kval32 = f->cfg.tangent_precision == 0 ? nullptr : f->d_kval32;
  return A;
}

void launch_scatter(pm_fem* f) {
  const int64_t nw = f->nkwave;
  if (nw <= 0) return;
  const unsigned grid = (unsigned)((nw + TPB / 64 - 1) / (TPB / 64));
  const bool lowp = f->cfg.tangent_precision != 0;
  if (!lowp) {
    const ScatWArgs<double> A = scatw_args(f, (const double*)f->d_ke);
    if (f->bs == 2) hipLaunchKernelGGL((k_scatter_w<2, double>), dim3(grid), dim3(TPB), 0, f->stream, A);
    else hipLaunchKernelGGL((k_scatter_w<3, double>), dim3(grid), dim3(TPB), 0, f->stream, A);
  } else {
    const ScatWArgs<float> A = scatw_args(f, (const float*)f->d_ke32);
    if (f->bs == 2) hipLaunchKernelGGL((k_scatter_w<2, float>), dim3(grid), dim3(TPB), 0, f->stream, A);
    else hipLaunchKernelGGL((k_scatter_w<3, float>), dim3(grid), dim3(TPB), 0, f->stream, A);
  }
}

void launch_residual_gather(pm_fem* f) {
  const int64_t n = f->npoint * f->ndim;
  const unsigned grid = (unsigned)((n + TPB - 1) / TPB);
  const RgArgs A = rg_args(f);
  if (f->ndim == 2) hipLaunchKernelGGL((k_rgather<2>), dim3(grid), dim3(TPB), 0, f->stream, A);
  else hipLaunchKernelGGL((k_rgather<3>), dim3(grid), dim3(TPB), 0, f->stream, A);
}

// ---- element passes of several handles in shared launches (serving) ---------------------------------
// Every kernel class (kernel, element kind and rule, work split, precision) of the requested passes becomes
// one launch over the union of the handles' groups; the jobs' descriptors and workgroup maps are staged in
// one device buffer.  Order: all updates, all residual gathers, all tangents, all K gathers.
namespace {

template <typename A>
struct Cls {
  std::vector<A> args;
  std::vector<int32_t> map;
  size_t shm = 0;
  size_t off_args = 0, off_map = 0;  // offsets in the staged buffer
  bool add(const A& a, int64_t nblk) {  // false when the job or its workgroups do not fit a map entry
    const int32_t j = (int32_t)args.size();
    if (j >= (1 << 11) || nblk >= (1 << 20)) return false;
    args.push_back(a);
    for (int64_t w = 0; w < nblk; w++) map.push_back((j << 20) | (int32_t)w);
    return true;
  }
};

inline size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

}  // namespace

int pass_batch_run(PassBatchBuf* buf, hipStream_t stream, std::vector<PassJob>& jobs) {
  using UK = std::tuple<int, int, int>;                  // kind, nip, epb
  using TK = std::tuple<int, int, int, int>;             // kind, nip, ch, shm
  std::map<UK, Cls<UpdArgs>> upd;
  std::map<int, Cls<RgArgs>> rg;
  std::map<TK, Cls<TanArgs<double>>> tan64;
  std::map<TK, Cls<TanArgs<float>>> tan32;
  std::map<int, Cls<ScatArgs<double>>> sc64;
  std::map<int, Cls<ScatArgs<float>>> sc32;
  std::vector<std::pair<pm_fem*, Group*>> mfma;           // groups whose tangent runs on the matrix cores
  std::vector<std::pair<pm_fem*, Group*>> h20;            // groups whose tangent runs as the persistent Hex20 kernel
  for (PassJob& J : jobs) {
    pm_fem* f = J.f;
    for (auto& g : f->groups) {
      if (g.nel == 0) continue;
      if (J.update && J.do_update && g.general_plasticity) { refuse(f, kGpUpdate); continue; }
      if (J.tangent && g.general_plasticity) { refuse(f, kGpStiffness); continue; }
      if (J.update) {
        const int epb = epb_update(g.ndim, g.nnode, g.nip);
        auto& c = upd[UK{g.kind, g.nip, epb}];
        c.shm = shm_update(g.ndim, g.nnode, g.nip, epb);
        if (!c.add(upd_args(f, g, J.do_update, epb), (g.nel + epb - 1) / epb)) return 1;
      }
      if (J.tangent) {
        if (tangent_kernel_of(f, g) == 2 || tangent_kernel_of(f, g) == 5) { mfma.push_back({f, &g}); continue; }
        if (runs_hex20(f, g)) { h20.push_back({f, &g}); continue; }
        if (!ke_range_ok(g.ke_base, g.nel, g.kew, packed_ke_values(g.nnode, g.ndim), f->n_ke_values)) { f->launch_refused = true; continue; }
        const int ch = tangent_split(f, g), epw = tangent_epw(g.nnode);
        const int kb = f->cfg.tangent_precision == 0 ? 8 : 4;
        const size_t shm = shm_tangent(g.ndim, g.nnode, g.nip, ch, kb);
        const TK key{g.kind, g.nip, ch, (int)shm};
        const int64_t nblk = (g.nel + epw - 1) / epw;
        if (kb == 8) { auto& c = tan64[key]; c.shm = shm; if (!c.add(tan_args(f, g, ch, f->d_ke, shm), nblk)) return 1; }
        else { auto& c = tan32[key]; c.shm = shm; if (!c.add(tan_args(f, g, ch, f->d_ke32, shm), nblk)) return 1; }
      }
    }
    if (J.residual && !rg[f->ndim].add(rg_args(f), (f->npoint * f->ndim + TPB - 1) / TPB)) return 1;
    if (J.tangent) {
      const int64_t nblk = (f->nnzu + TPB - 1) / TPB;
      if (f->cfg.tangent_precision == 0 ? !sc64[f->bs].add(scat_args(f, (const double*)f->d_ke), nblk)
                                        : !sc32[f->bs].add(scat_args(f, (const float*)f->d_ke32), nblk)) return 1;
    }
  }
  // stage every class's descriptors and map in one upload
  size_t total = 0;
  auto place = [&](auto& classes) {
    for (auto& kv : classes) {
      auto& c = kv.second;
      c.off_args = total; total += align256(c.args.size() * sizeof(c.args[0]));
      c.off_map = total; total += align256(c.map.size() * sizeof(int32_t));
    }
  };
  place(upd); place(rg); place(tan64); place(tan32); place(sc64); place(sc32);
  if (total > buf->cap) {
    if (buf->d) { (void)hipStreamSynchronize(stream); (void)hipFree(buf->d); buf->d = nullptr; buf->cap = 0; }
    if (hipMalloc(&buf->d, total) != hipSuccess) { buf->d = nullptr; return 1; }
    buf->cap = total;
  }
  buf->host.resize(total);
  auto copy_in = [&](auto& classes) {
    for (auto& kv : classes) {
      auto& c = kv.second;
      std::memcpy(buf->host.data() + c.off_args, c.args.data(), c.args.size() * sizeof(c.args[0]));
      std::memcpy(buf->host.data() + c.off_map, c.map.data(), c.map.size() * sizeof(int32_t));
    }
  };
  copy_in(upd); copy_in(rg); copy_in(tan64); copy_in(tan32); copy_in(sc64); copy_in(sc32);
  char* d = static_cast<char*>(buf->d);
  if (total && hipMemcpyAsync(d, buf->host.data(), total, hipMemcpyHostToDevice, stream) != hipSuccess) return 1;
  auto dargs = [&](auto& c) { return reinterpret_cast<decltype(c.args.data())>(d + c.off_args); };
  auto dmap = [&](auto& c) { return reinterpret_cast<const int32_t*>(d + c.off_map); };
  for (auto& kv : upd) {
    auto& c = kv.second;
    PM_DISPATCH_KIND(std::get<0>(kv.first), ND, NN,
                     hipLaunchKernelGGL((kb_update<ND, NN>), dim3((unsigned)c.map.size()), dim3(TPB), c.shm, stream, dargs(c), dmap(c)));
  }
  for (auto& kv : rg) {
    auto& c = kv.second;
    if (kv.first == 2) hipLaunchKernelGGL((kb_rgather<2>), dim3((unsigned)c.map.size()), dim3(TPB), 0, stream, dargs(c), dmap(c));
    else hipLaunchKernelGGL((kb_rgather<3>), dim3((unsigned)c.map.size()), dim3(TPB), 0, stream, dargs(c), dmap(c));
  }
  for (auto& kv : tan64) {
    auto& c = kv.second;
    PM_DISPATCH_KIND(std::get<0>(kv.first), ND, NN,
                     hipLaunchKernelGGL((kb_tangent<ND, NN, double>), dim3((unsigned)c.map.size()), dim3(TW), c.shm, stream, dargs(c), dmap(c)));
  }
  for (auto& kv : tan32) {
    auto& c = kv.second;
    PM_DISPATCH_KIND(std::get<0>(kv.first), ND, NN,
                     hipLaunchKernelGGL((kb_tangent<ND, NN, float>), dim3((unsigned)c.map.size()), dim3(TW), c.shm, stream, dargs(c), dmap(c)));
  }
  for (auto& mg : mfma) {  // one launch per group, on the batch's stream (the handle's owner waits for the batch)
    hipStream_t own = mg.first->stream;
    mg.first->stream = stream;
    (void)launch_matrix_core(mg.first, *mg.second);
    mg.first->stream = own;
  }
  for (auto& hg : h20) {  // likewise
    hipStream_t own = hg.first->stream;
    hg.first->stream = stream;
    (void)launch_tangent_hex20(hg.first, *hg.second);
    hg.first->stream = own;
  }
  for (auto& kv : sc64) {
    auto& c = kv.second;
    if (kv.first == 2) hipLaunchKernelGGL((kb_scatter<2, double>), dim3((unsigned)c.map.size()), dim3(TPB), 0, stream, dargs(c), dmap(c));
    else hipLaunchKernelGGL((kb_scatter<3, double>), dim3((unsigned)c.map.size()), dim3(TPB), 0, stream, dargs(c), dmap(c));
  }
  for (auto& kv : sc32) {
    auto& c = kv.second;
    if (kv.first == 2) hipLaunchKernelGGL((kb_scatter<2, float>), dim3((unsigned)c.map.size()), dim3(TPB), 0, stream, dargs(c), dmap(c));
    else hipLaunchKernelGGL((kb_scatter<3, float>), dim3((unsigned)c.map.size()), dim3(TPB), 0, stream, dargs(c), dmap(c));
  }
  if (hipGetLastError() != hipSuccess) return 1;
  return hipStreamSynchronize(stream) == hipSuccess ? 0 : 1;  // the staged descriptors stay valid until here
}

void launch_init_state(pm_fem* f, Group& g) {
  const int64_t n = g.nel * g.nip;
  if (n == 0) return;
  hipLaunchKernelGGL(k_init_state, dim3((unsigned)((n + TPB - 1) / TPB)), dim3(TPB), 0, f->stream, n, f->nsig, g.d_state,
                     g.param.z_ini, g.param.model == 1 ? 1 : 0);
}

void launch_geometry(pm_fem* f, Group& g) {
  if (g.nel == 0) return;
  const int64_t n = g.nel * g.nip;
  const unsigned grid = (unsigned)((n + TPB - 1) / TPB);
  PM_DISPATCH_KIND(g.kind, ND, NN, hipLaunchKernelGGL((k_geometry<ND, NN>), dim3(grid), dim3(TPB), 0, f->stream, g.nel, g.nip,
                                                      g.d_conn, f->d_X, g.d_tab, f->axisym ? 1 : 0, g.d_geo));
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

void reset_device_error(pm_fem* f) {
  const int32_t init[4] = {0, 0, -1, -1};  // no bits; key kErrKeyNone
  (void)hipMemcpyAsync(f->d_err, init, sizeof(init), hipMemcpyHostToDevice, f->stream);
  (void)hipStreamSynchronize(f->stream);
}

int check_device_error(pm_fem* f) {
  if (f->launch_refused) {
    f->launch_refused = false;
    f->err = f->refuse_msg ? f->refuse_msg : "internal error: an element-matrix range did not fit its buffer (kernel not launched)";
    f->refuse_msg = nullptr;
    return 1;
  }
  int32_t w[4] = {0, 0, 0, 0};
  PM_HIP(hipMemcpyAsync(w, f->d_err, sizeof(w), hipMemcpyDeviceToHost, f->stream));
  PM_HIP(hipStreamSynchronize(f->stream));
  const int32_t e = w[0];
  if (e == 0) return 0;
  uint64_t key = 0;
  std::memcpy(&key, w + 2, sizeof(key));
  reset_device_error(f);
  // the first failing IP: its cell and Gauss point (pmsim prints them before returning the error)
  f->err_cell = -1;
  f->err_gauss = -1;
  const uint64_t gi = key >> kErrKeyShift, ip = key & ((1ull << kErrKeyShift) - 1);
  if (key != kErrKeyNone && gi < f->groups.size()) {
    const Group& g = f->groups[gi];
    const uint64_t el = ip / (uint64_t)g.nip;
    if (el < (uint64_t)g.nel) { f->err_cell = g.cells[el]; f->err_gauss = (int32_t)(ip - el * g.nip); }
  }
  if (f->err_cell >= 0) std::printf("cell_id = %lld, gauss_id = %d\n", (long long)f->err_cell, f->err_gauss);
  if (e & ERR_DETJ) f->err = "negative or zero Jacobian determinant";
  else if (e & ERR_SIGD) f->err = "von Mises plastic update must not lead to zero sigma_d (sigma_d < 1e-10)";
  else if (e & ERR_NORMS) f->err = "von Mises stiffness cannot be computed with zero deviatoric norm (norm_s < 1e-10)";
  else f->err = "local matrix is not symmetric";
  return 1;
}

size_t update_shared_bytes(const Group& g) { return shm_update(g.ndim, g.nnode, g.nip, epb_update(g.ndim, g.nnode, g.nip)); }

}  // namespace pm
