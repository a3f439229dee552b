// Persistent matrix-core tangent for Hex20 elements with 27 integration points: FP32 operands (16x16x4
// MFMA) or BF16 operands (16x16x16), FP32 accumulation, element matrices stored in FP32 (the FP32 and BF16
// inexact-Newton tangents).
//
// For D = alpha I + beta I(x)I + gamma s(x)s (von_mises.rs:124-167) the element matrix is
//   K_ab = M_ab^T + tr(M_ab) 1 + C_ab,
//   M = sum_q (c alpha/2)_q g_q g_q^T,   C = sum_q (c beta)_q g_q g_q^T + (c gamma)_q ts_q ts_q^T,
// with g_q[(a,i)] = G_q[a][i] and ts_q = B_q^T s_q (the decomposition of the per-group matrix-core kernel).
// Each Gram is formed from a single row kind: the rows are scaled by the square roots of their weights
// (alpha > 0, beta >= 0 for nu >= 0), and the gamma term, whose weight is never positive, is subtracted
// by negating its B fragments in registers.  Node columns are padded to four (80 columns: 5x5 tiles of
// 16x16, 15 upper tiles), so that no 3x3 node block crosses a tile: the block transposes of M and its
// traces stay inside a tile, between the four lanes of a quad (DPP).  The IPs are taken in chunks (16 for
// BF16, the MFMA depth; 8 for FP32, two MFMA depths), the rows of a chunk staged in LDS in the operand type.
//
// Organisation as the persistent vector kernel (pm_tangent_hex20.hip): one workgroup per CU of up to eight
// wavefronts, one element per wavefront at a time from work counters, the shape-function derivatives in
// LDS once per workgroup, the next element's geometry, state and node ranks copied into the wavefront's LDS
// by direct global->LDS loads while the current one is processed; nodes in rank order, the matrix staged
// in LDS and written with 16-byte stores.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "pm_internal.hpp"

namespace pm {

namespace {

constexpr double ISQ2_M = 0.70710678118654752440;

constexpr int M_NIP = 27, M_NN = 20, M_ND = 3, M_NS = 6;
constexpr int M_NC = 80;                 // columns: 20 nodes x 4 (three dofs and a zero)
constexpr int M_T = 5;                   // 16-column tiles per dimension
constexpr int M_GW = 11;                 // per IP in the geometry array: J^-1 (9), det J, radius
constexpr int M_GEO_DW = 640, M_ST_DW = 448, M_RK_DW = 64;  // LDS copies of one element (dwords)
constexpr int M_RAW = (M_GEO_DW + M_ST_DW + M_RK_DW) / 2;    // doubles
constexpr int M_EW = packed_ke_values(M_NN, M_ND);           // 1892 values per element matrix
constexpr int M_STG = M_EW * 4;                              // output staging (FP32)
constexpr int M_PW = 20;                 // floats per IP in the row-factor table: A (9), T (9), b/a ratio, zero
constexpr int M_QP = 32;                 // IPs in the row-factor table (IPs 27..31: zero rows for the padded chunk)
constexpr int M_TAB = M_NIP * M_NN * 2 + M_NIP + 1;          // dN (float4 per IP and node), w, padding (doubles)
constexpr int M_WMAX = 8;

typedef __attribute__((address_space(3))) void lds_void_m;
typedef __attribute__((address_space(1))) int32_t gint_m;
typedef float f4_m __attribute__((ext_vector_type(4)));
typedef short s4_m __attribute__((ext_vector_type(4)));
typedef float V4 __attribute__((ext_vector_type(4)));
typedef float f2_m __attribute__((ext_vector_type(2)));
typedef __bf16 bf2_m __attribute__((ext_vector_type(2)));  // round to nearest even (v_cvt_pk_bf16_f32)

// operand precision P: 1 FP32 (v_mfma_f32_16x16x4_f32), 2 BF16 (v_mfma_f32_16x16x16_bf16)
template <int P> struct MOp;
template <> struct MOp<1> {
  typedef float E;                     // row element in LDS
  typedef float Frag;                  // one lane's MFMA operand: one k
  static constexpr int KS = 4;         // IPs per MFMA
  static constexpr int KC = 8;         // IPs per chunk
  static constexpr int RS = 9;         // elements per column row (8 used; odd: conflict-free single-k reads)
};
template <> struct MOp<2> {
  typedef uint16_t E;
  typedef s4_m Frag;                   // four consecutive k
  static constexpr int KS = 16;
  static constexpr int KC = 16;
  static constexpr int RS = 20;        // 16 used; 40 B: conflict-free 8-byte reads
};
template <int P> struct MShp {
  static constexpr int NCHK = (M_NIP + MOp<P>::KC - 1) / MOp<P>::KC;   // chunks (IPs past 27 are zero rows)
  static constexpr int KIND = M_NC * MOp<P>::RS;                       // elements per row kind
  static constexpr int ROWB = 3 * KIND * (int)sizeof(typename MOp<P>::E);  // bytes of a chunk's rows
  static constexpr int UB = ((ROWB > M_STG ? ROWB : M_STG) + 15) / 16 * 16;  // union bytes
  static constexpr int WAVE = M_RAW + M_QP * M_PW / 2 + UB / 8;        // doubles per wavefront (even)
  static_assert(WAVE % 2 == 0 && M_TAB % 2 == 0, "wavefront regions start 16-byte aligned");
  static_assert(MOp<P>::KC % MOp<P>::KS == 0 && MOp<P>::KC % 2 == 0, "chunks of whole MFMA depths and IP pairs");
};

struct MMat {
  int model;
  double K, G, H;
};

struct MArgs {
  int64_t nel;
  const uint8_t* rank; const double* geo; const double* tab; const double* st; float* ke;
  MMat mp; double thickness; int32_t* err; uint64_t ekey;
  int32_t* work;            // work counters (zero at launch)
  int ncnt;                 // counters in use (at most the number of workgroups)
};

__device__ __forceinline__ void lds_wait_m() { asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory"); }
__device__ __forceinline__ void order_m() { asm volatile("" ::: "memory"); }

// value of lane (p ^ S) of this lane's quad (DPP quad permutation)
template <int S>
__device__ __forceinline__ float quad_xor(float v) {
  constexpr int ctrl = S == 1 ? 0xB1 : S == 2 ? 0x4E : 0x1B;  // [1,0,3,2], [2,3,0,1], [3,2,1,0]
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), ctrl, 0xf, 0xf, false));
}
__device__ __forceinline__ float pick4(const float* x, int i) {
  return i == 0 ? x[0] : i == 1 ? x[1] : i == 2 ? x[2] : x[3];
}

template <int P>
__global__ __launch_bounds__(64 * M_WMAX) void k_tangent_hex20_mfma(MArgs A) {
  using O = MOp<P>;
  using S = MShp<P>;
  using E = typename O::E;
  using Frag = typename O::Frag;
  constexpr int KC = O::KC, RS = O::RS, KIND = S::KIND;
  constexpr int ITEMS = (KC / 2) * M_NN;            // staging items per chunk: (IP pair, node)
  constexpr int SR = (ITEMS + 63) / 64;             // staging rounds per chunk
  extern __shared__ __attribute__((aligned(16))) double sh[];
  const int64_t nel = A.nel;
  const int64_t nipg = nel * M_NIP;
  const double* __restrict__ geo = A.geo;
  const double* __restrict__ st = A.st;
  const uint8_t* __restrict__ rank = A.rank;
  float* __restrict__ ke = A.ke;
  const MMat mp = A.mp;
  const double thickness = A.thickness;
  int32_t* err = A.err;
  const uint64_t ekey = A.ekey;
  const int tid = threadIdx.x;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  V4* sdN = reinterpret_cast<V4*>(sh);                // [27][20]: dN/dxi of the node (three and a zero)
  double* sw = sh + M_NIP * M_NN * 2;
  {
    const double* __restrict__ tab = A.tab;
    for (int i = tid; i < M_NIP * M_NN; i += blockDim.x) {
      const double* d = tab + M_NIP + M_NIP * M_NN + i * M_ND;
      sdN[i] = V4{(float)d[0], (float)d[1], (float)d[2], 0.0f};
    }
    for (int i = tid; i < M_NIP; i += blockDim.x) sw[i] = tab[i];
  }
  const int ncnt = A.ncnt, cidx = (int)(blockIdx.x % (unsigned)ncnt);

  double* raw = sh + M_TAB + (size_t)wv * S::WAVE;
  double* rgeo = raw;
  double* rst = raw + M_GEO_DW / 2;
  double* rrk = rst + M_ST_DW / 2;
  const uint8_t* srk = reinterpret_cast<const uint8_t*>(rrk);
  float* sP = reinterpret_cast<float*>(raw + M_RAW);                   // [32][20] row factors per IP
  char* uni = reinterpret_cast<char*>(sP + M_QP * M_PW);               // rows of a chunk | output staging
  E* rows = reinterpret_cast<E*>(uni);                                 // [3 kinds][80 columns][RS]
  float* stage = reinterpret_cast<float*>(uni);
  {
    const int lane = tid & 63;
    for (int i = M_NIP * M_PW + lane; i < M_QP * M_PW; i += 64) sP[i] = 0.0f;  // padding IPs: zero rows
  }
  __syncthreads();  // the only workgroup-wide wait

  int32_t* work = A.work + cidx * kWorkStride;
  auto ticket = [&]() -> int32_t {
    int32_t t = 0;
    gint_m* wp = (gint_m*)work;
    asm volatile("" : "+v"(wp));
    if ((threadIdx.x & 63) == 0) t = __hip_atomic_fetch_add(wp, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return t;
  };
  // the next element's state (sigma, lambda, elastic flag: one component of 27 doubles per load), geometry and
  // node ranks, copied into this wavefront's LDS: uniform bases, lane offsets of four bytes
  auto prefetch = [&](int64_t e) {
    int lane = tid & 63;
    asm volatile("" : "+v"(lane));
    const uint32_t lo = 4u * (uint32_t)lane;
    if (lane < 2 * M_NIP) {
#pragma unroll
      for (int c = 0; c < M_NS + 2; c++) {
        const int cg = c < M_NS ? c : c + 1;
        const char* src = reinterpret_cast<const char*>(st + cg * nipg + e * M_NIP);
        __builtin_amdgcn_global_load_lds(reinterpret_cast<const uint32_t*>(src + lo), (lds_void_m*)(rst + M_NIP * c), 4, 0, 0);
      }
    }
    const char* gsrc = reinterpret_cast<const char*>(geo + e * (M_NIP * M_GW));
#pragma unroll
    for (int k = 0; k < M_GEO_DW / 64 - 1; k++)
      __builtin_amdgcn_global_load_lds(reinterpret_cast<const uint32_t*>(gsrc + lo + 256 * k), (lds_void_m*)(rgeo + 32 * k), 4, 0, 0);
    if (lane < M_NIP * M_GW * 2 - (M_GEO_DW - 64))
      __builtin_amdgcn_global_load_lds(reinterpret_cast<const uint32_t*>(gsrc + lo + 4 * (M_GEO_DW - 64)), (lds_void_m*)(rgeo + M_GEO_DW / 2 - 32), 4, 0, 0);
    if (lane < M_NN / 4)
      __builtin_amdgcn_global_load_lds(reinterpret_cast<const uint32_t*>(rank + e * M_NN) + lane, (lds_void_m*)rrk, 4, 0, 0);
    order_m();
  };
  int32_t tk = ticket();
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  int64_t cur = cidx + (int64_t)ncnt * __builtin_amdgcn_readfirstlane(tk);
  if (cur >= nel) return;  // uniform over the wavefront
  prefetch(cur);
  tk = ticket();
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  int64_t nxt = cidx + (int64_t)ncnt * __builtin_amdgcn_readfirstlane(tk);

  // epilogue lane roles: column j = 4 bp + p of a tile, rows 4 g .. 4 g + 3
  int lane0 = tid & 63;
  asm volatile("" : "+v"(lane0));
  const int p = lane0 & 3;
  const bool p0 = (p & 1) != 0, p1 = (p & 2) != 0;
  const float d0 = p == 0 ? 1.0f : 0.0f, d1 = p == 1 ? 1.0f : 0.0f, d2 = p == 2 ? 1.0f : 0.0f;

  while (cur < nel) {
    const int64_t e = cur;
    int lane = tid & 63;
    asm volatile("" : "+v"(lane));
    // per IP, one lane per IP: D coefficients (von_mises.rs:124-167), then the row factors of the IP:
    // A = sqrt(c alpha/2) J^-T, T = sqrt(-c gamma) s J^-T (s the deviator as a symmetric 3x3 matrix) and
    // sqrt(c beta) / sqrt(c alpha/2), so that the rows of node m are A dN_m, ratio * A dN_m and T dN_m
    if (lane < M_NIP) {
      const int q = lane;
      const int64_t ip = e * M_NIP + q;
      const double det = rgeo[q * M_GW + 9];
      if (!(det > 0.0)) raise_error(err, ERR_DETJ, ekey + (uint64_t)ip);
      const double coef = det * sw[q] * thickness;
      double alpha, beta, gamma = 0.0;
      double s[6] = {0, 0, 0, 0, 0, 0};
      if (mp.model == 0 || rst[7 * M_NIP + q] != 0.0) {
        alpha = 2.0 * mp.G;
        beta = mp.K - alpha / 3.0;
      } else {
        double sig[M_NS];
#pragma unroll
        for (int c = 0; c < M_NS; c++) sig[c] = rst[c * M_NIP + q];
        const double lam = rst[6 * M_NIP + q];
        const double sm_ = (sig[0] + sig[1] + sig[2]) / 3.0;
        double ss = 0.0;
#pragma unroll
        for (int c = 0; c < M_NS; c++) { s[c] = sig[c] - (c < 3 ? sm_ : 0.0); ss += s[c] * s[c]; }
        const double sd = sqrt(1.5 * ss);
        const double sdtr = sd + lam * 3.0 * mp.G;
        const double norm_s = sd * 0.81649658092772603273;  // sqrt(2/3)
        if (norm_s < 1e-10) raise_error(err, ERR_NORMS, ekey + (uint64_t)ip);
        const double d = 3.0 * mp.G + mp.H;
        alpha = 2.0 * mp.G * (1.0 - lam * 3.0 * mp.G / sdtr);
        gamma = 6.0 * mp.G * mp.G * (lam / sdtr - 1.0 / d) / (norm_s * norm_s);
        beta = mp.K - alpha / 3.0;
      }
      const float wa = sqrtf((float)(coef * alpha * 0.5));
      const float wb = sqrtf(fmaxf((float)(coef * beta), 0.0f));
      const float wg = sqrtf(fmaxf((float)(-coef * gamma), 0.0f));
      float Ji[9];
#pragma unroll
      for (int k = 0; k < 9; k++) Ji[k] = (float)rgeo[q * M_GW + k];
      const float sxy = (float)(s[3] * ISQ2_M), syz = (float)(s[4] * ISQ2_M), sxz = (float)(s[5] * ISQ2_M);
      const float Sm[3][3] = {{(float)s[0], sxy, sxz}, {sxy, (float)s[1], syz}, {sxz, syz, (float)s[2]}};
      float* Pq = sP + q * M_PW;
      float out[M_PW];
#pragma unroll
      for (int i = 0; i < 3; i++)
#pragma unroll
        for (int k = 0; k < 3; k++) {
          out[3 * i + k] = wa * Ji[3 * k + i];
          out[9 + 3 * i + k] = wg * (Sm[i][0] * Ji[3 * k] + Sm[i][1] * Ji[3 * k + 1] + Sm[i][2] * Ji[3 * k + 2]);
        }
      out[18] = wa > 0.0f ? wb / wa : 0.0f;
      out[19] = 0.0f;
#pragma unroll
      for (int k = 0; k < M_PW / 4; k++)
        reinterpret_cast<V4*>(Pq)[k] = V4{out[4 * k], out[4 * k + 1], out[4 * k + 2], out[4 * k + 3]};
    }
    // column of each staging round's node (the node ranks are read before the next element's replace them)
    int colr[SR];
#pragma unroll
    for (int r = 0; r < SR; r++) {
      const int w = min(lane + 64 * r, ITEMS - 1);
      colr[r] = 4 * srk[w % M_NN];
    }
    int32_t tk2 = ticket();  // the element after nxt, read at the end of this element
    const bool next = nxt < nel;  // uniform
    lds_wait_m();  // this element's state, geometry and ranks have been read; the row factors are written
    if (next) prefetch(nxt);
    f4_m accM[15], accC[15];
    // Gram products of a chunk on the upper tiles: A fragment = 16 columns of tile I x the k of this lane's
    // group (four consecutive IPs for BF16, one for FP32), B fragment likewise for tile J; the gamma Gram is
    // subtracted by negating its A fragments.  The first products of an element start from zero.
    auto grams = [&](auto first) __attribute__((always_inline)) {
#pragma unroll
      for (int ks = 0; ks < KC / O::KS; ks++) {
        const int fo = (lane & 15) * RS + (P == 2 ? 4 * (lane >> 4) : O::KS * ks + (lane >> 4));
        const bool z = decltype(first)::value && ks == 0;
        int t = 0;
#pragma unroll
        for (int I = 0; I < M_T; I++) {
          const Frag a1 = *reinterpret_cast<const Frag*>(rows + 0 * KIND + 16 * I * RS + fo);
          const Frag a2 = *reinterpret_cast<const Frag*>(rows + 1 * KIND + 16 * I * RS + fo);
          Frag a3 = *reinterpret_cast<const Frag*>(rows + 2 * KIND + 16 * I * RS + fo);
          if constexpr (P == 2) a3 ^= (short)0x8000;  // -sqrt(-c gamma) ts
          else a3 = -a3;
#pragma unroll
          for (int J = I; J < M_T; J++, t++) {
            const Frag b1 = *reinterpret_cast<const Frag*>(rows + 0 * KIND + 16 * J * RS + fo);
            const Frag b2 = *reinterpret_cast<const Frag*>(rows + 1 * KIND + 16 * J * RS + fo);
            const Frag b3 = *reinterpret_cast<const Frag*>(rows + 2 * KIND + 16 * J * RS + fo);
            const f4_m zero = {0.0f, 0.0f, 0.0f, 0.0f};
            if constexpr (P == 2) {
              accM[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a1, b1, z ? zero : accM[t], 0, 0, 0);
              accC[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a2, b2, z ? zero : accC[t], 0, 0, 0);
              accC[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a3, b3, accC[t], 0, 0, 0);
            } else {
              accM[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b1, z ? zero : accM[t], 0, 0, 0);
              accC[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(a2, b2, z ? zero : accC[t], 0, 0, 0);
              accC[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(a3, b3, accC[t], 0, 0, 0);
            }
          }
        }
      }
    };
#pragma unroll 1
    for (int h = 0; h < S::NCHK; h++) {
      if (h > 0) lds_wait_m();  // the previous chunk's rows consumed
      // rows of IPs KC h + 2 pr, KC h + 2 pr + 1 for node m (item pr * 20 + m): three kinds x three dofs, the
      // fourth column of the node zero; IPs past 27 have zero row factors
#pragma unroll
      for (int r = 0; r < SR; r++) {
        const int w = lane + 64 * r;
        if (ITEMS % 64 == 0 || r < SR - 1 || w < ITEMS) {
          const int pr = w / M_NN, m = w - pr * M_NN;
          float v[2][9];
#pragma unroll
          for (int t = 0; t < 2; t++) {
            const int q = KC * h + 2 * pr + t;
            const V4 dn = sdN[min(q, M_NIP - 1) * M_NN + m];
            const V4* Pq = reinterpret_cast<const V4*>(sP + q * M_PW);
            const V4 f0 = Pq[0], f1 = Pq[1], f2 = Pq[2], f3 = Pq[3], f4 = Pq[4];
            const float Pf[M_PW] = {f0[0], f0[1], f0[2], f0[3], f1[0], f1[1], f1[2], f1[3], f2[0], f2[1],
                                    f2[2], f2[3], f3[0], f3[1], f3[2], f3[3], f4[0], f4[1], f4[2], f4[3]};
#pragma unroll
            for (int i = 0; i < 3; i++) {
              const float va = fmaf(Pf[3 * i], dn[0], fmaf(Pf[3 * i + 1], dn[1], Pf[3 * i + 2] * dn[2]));
              v[t][i] = va;
              v[t][3 + i] = Pf[18] * va;
              v[t][6 + i] = fmaf(Pf[9 + 3 * i], dn[0], fmaf(Pf[9 + 3 * i + 1], dn[1], Pf[9 + 3 * i + 2] * dn[2]));
            }
          }
#pragma unroll
          for (int z = 0; z < 3; z++) {
            E* rr = rows + z * KIND + colr[r] * RS + 2 * pr;
#pragma unroll
            for (int i = 0; i < 3; i++) {
              if constexpr (P == 2) {
                const f2_m pv = {v[0][3 * z + i], v[1][3 * z + i]};
                *reinterpret_cast<uint32_t*>(rr + i * RS) = __builtin_bit_cast(uint32_t, __builtin_convertvector(pv, bf2_m));
              } else {
                rr[i * RS] = v[0][3 * z + i];
                rr[i * RS + 1] = v[1][3 * z + i];
              }
            }
            if constexpr (P == 2) *reinterpret_cast<uint32_t*>(rr + 3 * RS) = 0u;
            else { rr[3 * RS] = 0.0f; rr[3 * RS + 1] = 0.0f; }
          }
        }
      }
      lds_wait_m();
      if (h == 0) grams(std::true_type{});
      else grams(std::false_type{});
      order_m();
    }
    // epilogue per tile: K = M^T (per node block) + tr(M) 1 + C into the staging at the stored offsets of the
    // upper node blocks.  The four lanes of a quad hold one 4x4 node block of M by columns (lane p: column p);
    // a two-step butterfly over the quad (exchange lane bit 0 with register bit 0, then bit 1 with bit 1)
    // leaves row p in lane p, i.e. column p of M^T.  The padding row and column of a node are zero, so the
    // trace over four is the trace over three.
    lds_wait_m();  // the rows are consumed (the staging overlays them)
    {
      const int j = lane & 15, g = lane >> 4, bp = j >> 2;
      int t = 0;
#pragma unroll
      for (int I = 0; I < M_T; I++)
#pragma unroll
        for (int J = I; J < M_T; J++, t++) {
          const float x0 = accM[t][0], x1 = accM[t][1], x2 = accM[t][2], x3 = accM[t][3];
          float ra = quad_xor<1>(p0 ? x0 : x1), rb = quad_xor<1>(p0 ? x2 : x3);
          const float y0 = p0 ? ra : x0, y1 = p0 ? x1 : ra, y2 = p0 ? rb : x2, y3 = p0 ? x3 : rb;
          ra = quad_xor<2>(p1 ? y0 : y2);
          rb = quad_xor<2>(p1 ? y1 : y3);
          const float z0 = p1 ? ra : y0, z1 = p1 ? rb : y1, z2 = p1 ? y2 : ra;  // z_i = M(p, i)
          const float dg = fmaf(x2, d2, fmaf(x1, d1, x0 * d0));          // M(p, p)
          const float t1 = dg + quad_xor<1>(dg);
          const float tr = t1 + quad_xor<2>(t1);
          const int a = 4 * I + g, b = 4 * J + bp;
          if (p < 3 && (I < J || g <= bp)) {
            float* o = stage + ke_block_offset(M_NN, M_ND, a, b) + p;
            o[0] = fmaf(tr, d0, accC[t][0] + z0);
            o[3] = fmaf(tr, d1, accC[t][1] + z1);
            o[6] = fmaf(tr, d2, accC[t][2] + z2);
          }
        }
      if (lane == 0) { stage[ke_block_offset(M_NN, M_ND, 6, 6) - 1] = 0.0f; stage[M_EW - 1] = 0.0f; }
    }
    lds_wait_m();
    {
      constexpr int NPIECE = M_EW / 4, ROUNDS = (NPIECE + 63) / 64;  // 473 pieces, 8 rounds
      V4* dst = reinterpret_cast<V4*>(ke + e * M_EW);
      const V4* src = reinterpret_cast<const V4*>(stage);
#pragma unroll
      for (int r = 0; r < ROUNDS; r++) {
        const int i = min(lane + 64 * r, NPIECE - 1);  // surplus lanes repeat the last piece (same value)
        dst[i] = src[i];
      }
      static_assert(ROUNDS == 8, "the wait below counts eight stores");
    }
    order_m();
    // the next element's data and the ticket precede this element's eight stores; vector memory operations
    // complete in order
    asm volatile("s_waitcnt vmcnt(8)" ::: "memory");
    cur = nxt;
    asm volatile("" : "+v"(tk2));
    nxt = cidx + (int64_t)ncnt * __builtin_amdgcn_readfirstlane(tk2);
  }
}

__global__ void k_zero_work_m(int32_t* w) {
  if (threadIdx.x < kWorkCounters) w[threadIdx.x * kWorkStride] = 0;
}

template <int P>
size_t mfma20_lds(int nw) { return sizeof(double) * ((size_t)M_TAB + (size_t)nw * MShp<P>::WAVE); }

template <int P>
int mfma20_waves(const pm_fem* f) {
  int nw = M_WMAX;
  while (nw > 2 && (int64_t)mfma20_lds<P>(nw) > f->lds_per_block) nw--;
  return (int64_t)mfma20_lds<P>(nw) <= f->lds_per_block ? nw : 0;
}

}  // namespace

// true when a group's reduced-precision tangent can run as the persistent Hex20 matrix-core kernel: Hex20 cells
// with 27 IPs, an FP32 or BF16 tangent, a Poisson ratio >= 0 (beta >= 0, so that every Gram weight has a real
// square root), and a device whose workgroups may hold its LDS
bool tangent_hex20_mfma_fits(const pm_fem* f, const Group& g) {
  const int prec = f->cfg.tangent_precision;
  if (g.kind != HEX20 || g.nip != M_NIP || f->axisym || g.plane_stress || g.param.poisson < 0.0) return false;
  return prec == 1 ? mfma20_waves<1>(f) > 0 : prec == 2 ? mfma20_waves<2>(f) > 0 : false;
}

template <int P>
static void launch_hex20_mfma(pm_fem* f, Group& g) {
  const int nw = mfma20_waves<P>(f);
  const size_t shm = mfma20_lds<P>(nw);
  const int per_cu = std::max<int>(1, (int)(f->lds_per_cu / (int64_t)shm));
  const int64_t grid = std::min<int64_t>((int64_t)f->n_cu * per_cu, (g.nel + nw - 1) / nw);
  static thread_local bool attr_set = false;
  if (!attr_set) {
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_tangent_hex20_mfma<P>), hipFuncAttributeMaxDynamicSharedMemorySize, (int)f->lds_per_block);
    (void)hipGetLastError();
    attr_set = true;
  }
  MMat mp;
  mp.model = g.param.model; mp.K = g.K; mp.G = g.G; mp.H = g.param.hh;
  const uint64_t ekey = (uint64_t)(&g - f->groups.data()) << kErrKeyShift;
  hipLaunchKernelGGL(k_zero_work_m, dim3(1), dim3(64), 0, f->stream, f->d_work);  // a kernel, not a memset (graph replays)
  MArgs A{g.nel, g.d_rank, g.d_geo, g.d_tab, g.d_state, f->d_ke32 + g.ke_base, mp, f->thickness, f->d_err, ekey, f->d_work,
          (int)std::min<int64_t>(kWorkCounters, grid)};
  hipLaunchKernelGGL((k_tangent_hex20_mfma<P>), dim3((unsigned)grid), dim3(64 * nw), shm, f->stream, A);
}

// Launches the persistent Hex20 matrix-core tangent of one group; false when it does not fit the group
bool launch_tangent_hex20_mfma(pm_fem* f, Group& g) {
  if (!tangent_hex20_mfma_fits(f, g)) return false;
  if (g.nel == 0) return true;
  if (!ke_range_ok(g.ke_base, g.nel, g.kew, M_EW, f->n_ke_values)) {
    f->launch_refused = true;  // reported by check_device_error; never launched out of range
    return true;
  }
  if (f->cfg.tangent_precision == 1) launch_hex20_mfma<1>(f, g);
  else launch_hex20_mfma<2>(f, g);
  return true;
}

}  // namespace pm
