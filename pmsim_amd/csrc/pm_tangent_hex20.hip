// Persistent vector tangent for Hex20 elements with 27 integration points (FP64 products; the element
// matrices stored in FP64, or rounded to FP32 for the FP32 tangent).
//
// Same closed form as the node-block vector kernel of pm_kernels.hip (von_mises.rs:124-167,
// element_solid.rs:168-185):
//   K_ab = sum_q alpha/2 (Ga.Gb) 1 + alpha/2 Gb Ga^T + beta Ga Gb^T + gamma tsa tsb^T
// except that the elastic IPs (alpha = 2G, beta = lambda, gamma = 0) are summed first as one Gram
// Y_ab = sum_q c Ga Gb^T, turned once into G (Y_ab^T + tr(Y_ab) 1) + lambda Y_ab; the plastic IPs then add
// their terms in the node-block kernel's order (the element matrices agree with that kernel's to rounding).
// The organisation also differs:
//  * one workgroup per CU of up to twelve wavefronts; every wavefront works through elements one at a time,
//    each taken from a work counter one element ahead (so wavefronts that get more issue cycles take more
//    elements and all finish together; the counters are spread over many addresses), and never waits for
//    another wavefront;
//  * the shape-function derivatives of the 27 IPs sit in LDS once per workgroup;
//  * the IP geometry (J^-1, det J) and state of the next element are copied from HBM into the
//    wavefront's LDS by direct global->LDS loads while the last IPs of the current element are
//    accumulated and its matrix is written, so an element starts with its data already in LDS;
//  * node vectors are staged in the slots of the nodes' ranks (nodes sorted by global id), so the 55 active
//    lanes own fixed 2x2 tiles of upper node blocks in rank order, the stored order (ke_block_offset); the
//    matrix is staged through LDS in its two halves (rank rows 0-5, 6-19; one pass in FP32) and written with
//    16-byte stores.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "pm_internal.hpp"

namespace pm {

namespace {

constexpr double ISQ2_H = 0.70710678118654752440;

constexpr int H_NIP = 27, H_NN = 20, H_ND = 3, H_NS = 6;
constexpr int H_NH = 10, H_NT = 55;     // node pairs, upper 2x2 tiles of node blocks
constexpr int H_PW = 7;                 // node vector: G (3), unused, ts (3): odd stride
constexpr int H_CH = 3, H_NCH = 9;      // IPs per chunk (60 staging lanes), chunks
constexpr int H_QW = 19;                // per IP slot: coefficients (3), deviator (6), J^-1 (9), IP index
constexpr int H_GW = 11;                // per IP in the geometry array: J^-1 (9), det J, radius
constexpr int H_GEO_DW = 640;           // geometry of one element in LDS, dwords (594 used, 10 loads)
constexpr int H_ST_DW = 448;            // state of one element in LDS, dwords (8 x 27 doubles used, 7 loads)
constexpr int H_RK_DW = 64;             // ranks of one element's nodes in LDS, dwords (5 used, one load)
constexpr int H_RAW = (H_GEO_DW + H_ST_DW + H_RK_DW) / 2;  // doubles
constexpr int H_EW = packed_ke_values(H_NN, H_ND);     // 1892 values per element matrix
constexpr int H_HALF = H_EW / 2;                       // 946: one half of the stored matrix
constexpr int H_U = H_HALF;                            // doubles: per-IP data and node vectors, or output
constexpr int H_WAVE = H_RAW + H_U;                    // doubles per wavefront
constexpr int H_TAB = H_NIP * H_NN * H_ND + H_NIP + 1;  // dN, w, padding (doubles, even)
constexpr int H_WMAX = 12;
static_assert(H_NIP * H_QW + H_CH * H_NN * H_PW <= H_U, "per-IP data and node vectors fit the union");
static_assert(H_WAVE % 2 == 0 && H_TAB % 2 == 0, "wavefront regions start 16-byte aligned");
static_assert(ke_split_row(H_NN, H_ND) == 6 && ke_block_offset(H_NN, H_ND, 6, 6) == H_HALF, "halves of the stored matrix");

typedef __attribute__((address_space(3))) void lds_void;
typedef __attribute__((address_space(1))) int32_t gint;

struct H20Mat {
  int model;                // 0 linear elastic, 1 von Mises
  double K, G, H;
};

template <typename KT>
struct H20Args {
  int64_t nel;
  const uint8_t* rank; const double* geo; const double* tab; const double* st; KT* ke;
  H20Mat mp; double thickness; int32_t* err; uint64_t ekey;
  int32_t* work;            // work counters (zero at launch)
  int ncnt;                 // counters in use (at most the number of workgroups)
};

__device__ __forceinline__ void lds_wait() { asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory"); }
__device__ __forceinline__ void order() { asm volatile("" ::: "memory"); }

template <typename KT>
__global__ __launch_bounds__(64 * H_WMAX) void k_tangent_hex20(H20Args<KT> A) {
  extern __shared__ __attribute__((aligned(16))) double sh[];
  const int64_t nel = A.nel;
  const int64_t nipg = nel * H_NIP;
  const double* __restrict__ geo = A.geo;
  const double* __restrict__ st = A.st;
  KT* __restrict__ ke = A.ke;
  const H20Mat mp = A.mp;
  const double thickness = A.thickness;
  int32_t* err = A.err;
  const uint64_t ekey = A.ekey;
  const int tid = threadIdx.x;
  const int nw = blockDim.x >> 6;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  // shape-function derivatives and weights of the 27 IPs, once per workgroup
  double* sdN = sh;
  double* sw = sh + H_NIP * H_NN * H_ND;
  {
    const double* __restrict__ tab = A.tab;
    for (int i = tid; i < H_NIP * H_NN * H_ND; i += blockDim.x) sdN[i] = tab[H_NIP + H_NIP * H_NN + i];
    for (int i = tid; i < H_NIP; i += blockDim.x) sw[i] = tab[i];
  }
  __syncthreads();  // the only workgroup-wide wait
  // elements: this workgroup's counter c hands out c, c + ncnt, c + 2 ncnt, ... (a ticket each); every
  // wavefront ends once its element index passes nel
  const int ncnt = A.ncnt, cidx = (int)(blockIdx.x % (unsigned)ncnt);
  int32_t* work = A.work + cidx * kWorkStride;
  // a ticket: one lane's atomic add, read back (readfirstlane) only at the end of the element, when it has
  // long arrived.  The address passes through an empty asm so that the compiler does not turn the add into
  // a wavefront-wide one whose result it would wait for at once.
  auto ticket = [&]() -> int32_t {
    int32_t t = 0;
    gint* wp = (gint*)work;
    asm volatile("" : "+v"(wp));
    if ((threadIdx.x & 63) == 0) t = __hip_atomic_fetch_add(wp, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return t;
  };

  double* raw = sh + H_TAB + (size_t)wv * H_WAVE;   // [geometry: 27 x 11 | state: 8 x 27] of one element
  double* rgeo = raw;
  double* rst = raw + H_GEO_DW / 2;                 // sigma (6), lambda, elastic flag; 27 each
  double* rrk = rst + H_ST_DW / 2;                  // ranks of the 20 nodes (bytes)
  const uint8_t* srk = reinterpret_cast<const uint8_t*>(rrk);
  double* sQ = raw + H_RAW;                         // [27 slots][19] per-IP coefficients, deviator, J^-1
  double* sP = sQ + H_NIP * H_QW;                   // [3 IPs][20 nodes][7] node vectors of a chunk
  double* sU = sQ;                                  // output staging (over the per-IP data)

  // Direct global->LDS copies of element e's data into this wavefront's region, in one burst: the state
  // (components 0-5, 7, 8 of the state arrays: 54 dwords each, one load per component), the geometry (594
  // contiguous dwords in ten loads) and the ranks (5 dwords); uniform bases, lane offsets of four bytes.
  const uint8_t* __restrict__ rank = A.rank;
  auto prefetch = [&](int64_t e) {
    int lane = tid & 63;
    asm volatile("" : "+v"(lane));
    const uint32_t lo = 4u * (uint32_t)lane;
    if (lane < 2 * H_NIP) {
#pragma unroll
      for (int c = 0; c < H_NS + 2; c++) {
        const int cg = c < H_NS ? c : c + 1;  // z (component 6) is not needed
        const char* src = reinterpret_cast<const char*>(st + cg * nipg + e * H_NIP);
        __builtin_amdgcn_global_load_lds(reinterpret_cast<const uint32_t*>(src + lo), (lds_void*)(rst + H_NIP * c), 4, 0, 0);
      }
    }
    const char* gsrc = reinterpret_cast<const char*>(geo + e * (H_NIP * H_GW));
#pragma unroll
    for (int k = 0; k < H_GEO_DW / 64 - 1; k++)
      __builtin_amdgcn_global_load_lds(reinterpret_cast<const uint32_t*>(gsrc + lo + 256 * k), (lds_void*)(rgeo + 32 * k), 4, 0, 0);
    if (lane < H_NIP * H_GW * 2 - (H_GEO_DW - 64))
      __builtin_amdgcn_global_load_lds(reinterpret_cast<const uint32_t*>(gsrc + lo + 4 * (H_GEO_DW - 64)), (lds_void*)(rgeo + H_GEO_DW / 2 - 32), 4, 0, 0);
    if (lane < H_NN / 4)
      __builtin_amdgcn_global_load_lds(reinterpret_cast<const uint32_t*>(rank + e * H_NN) + lane, (lds_void*)rrk, 4, 0, 0);
    order();
  };
  int32_t tk = ticket();
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  int64_t cur = cidx + (int64_t)ncnt * __builtin_amdgcn_readfirstlane(tk);
  if (cur >= nel) return;  // uniform over the wavefront
  prefetch(cur);
  tk = ticket();
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  int64_t nxt = cidx + (int64_t)ncnt * __builtin_amdgcn_readfirstlane(tk);
  // elastic IPs: D = 2G I + lambda I(x)I, so their part of K_ab is G (Y_ab^T + tr(Y_ab) 1) + lambda Y_ab with
  // Y = sum_q c_q G_a G_b^T, one Gram; they are taken first and Y converted before the plastic IPs
  const double Gm = mp.G, lamb = mp.K - 2.0 * mp.G / 3.0;

  while (cur < nel) {
    const int64_t e = cur;
    // per-lane roles, re-derived per element from an opaque copy of the lane index (see prefetch)
    int lane = tid & 63;
    asm volatile("" : "+v"(lane));
    // this lane's tile (I <= J): node blocks (2I+u, 2J+v)
    const bool active = lane < H_NT;
    int I = 0, J = 0;
    {
      int rem = active ? lane : 0;
      while (rem >= H_NH - I) { rem -= H_NH - I; I++; }
      J = I + rem;
    }
    // this lane's node vector in a chunk's staging: IP ql of the chunk, node m
    const bool stager = lane < H_CH * H_NN;
    const int sql = lane / H_NN, sm = lane - sql * H_NN;
    // per IP: D coefficients (von_mises.rs:124-167) and deviator, one lane per IP, written to the IP's slot:
    // the elastic IPs first (in IP order), then the plastic ones
    const bool el = lane < H_NIP && (mp.model == 0 || rst[7 * H_NIP + (lane < H_NIP ? lane : 0)] != 0.0);
    const uint64_t emask = __ballot(el);
    const int n_el = __builtin_popcountll(emask);  // uniform
    if (lane < H_NIP) {
      const int q = lane;
      const uint64_t below = (1ull << q) - 1ull;
      const int slot = el ? __builtin_popcountll(emask & below) : n_el + __builtin_popcountll(~emask & below);
      const int64_t ip = e * H_NIP + q;
      const double det = rgeo[q * H_GW + 9];
      if (!(det > 0.0)) raise_error(err, ERR_DETJ, ekey + (uint64_t)ip);
      double coef = det * sw[q] * thickness * 1.0;
      double alpha, beta, gamma = 0.0;
      double s[6] = {0, 0, 0, 0, 0, 0};
      if (el) {
        alpha = 2.0 * mp.G;
        beta = mp.K - alpha / 3.0;
      } else {
        double sig[H_NS];
#pragma unroll
        for (int c = 0; c < H_NS; c++) sig[c] = rst[c * H_NIP + q];
        double lam = rst[6 * H_NIP + q];
        double sm_ = (sig[0] + sig[1] + sig[2]) / 3.0, ss = 0.0;
#pragma unroll
        for (int c = 0; c < H_NS; c++) { s[c] = sig[c] - (c < 3 ? sm_ : 0.0); ss += s[c] * s[c]; }
        double sd = sqrt(1.5 * ss);
        double sdtr = sd + lam * 3.0 * mp.G;
        double norm_s = sd * 0.81649658092772603273;  // sqrt(2/3)
        if (norm_s < 1e-10) raise_error(err, ERR_NORMS, ekey + (uint64_t)ip);
        double d = 3.0 * mp.G + mp.H;
        alpha = 2.0 * mp.G * (1.0 - lam * 3.0 * mp.G / sdtr);
        gamma = 6.0 * mp.G * mp.G * (lam / sdtr - 1.0 / d) / (norm_s * norm_s);
        beta = mp.K - alpha / 3.0;
      }
      double* Q = sQ + slot * H_QW;
      Q[0] = el ? coef : coef * alpha * 0.5;  // an elastic slot keeps c (its Gram weight)
      Q[1] = coef * beta;
      Q[2] = coef * gamma;
#pragma unroll
      for (int c = 0; c < 6; c++) Q[3 + c] = s[c];
#pragma unroll
      for (int k = 0; k < 9; k++) Q[9 + k] = rgeo[q * H_GW + k];
      Q[18] = (double)q;
    }
    const int myrank = srk[lane % H_NN];  // rank of this lane's staging node
    tk = ticket();  // the element after nxt, read at the end of this element (it has long arrived)
    const bool next = nxt < nel;  // uniform
    lds_wait();  // this element's state, geometry and ranks have been read (and its slots written)
    if (next) prefetch(nxt);
    double K[2][2][3][3];
    double Dg[2][2];
#pragma unroll
    for (int u = 0; u < 2; u++)
#pragma unroll
      for (int v = 0; v < 2; v++) {
        Dg[u][v] = 0.0;
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
          for (int j = 0; j < 3; j++) K[u][v][i][j] = 0.0;
      }
    // the elastic IPs' Gram Y (held in K) to their part of K: G Y^T + lambda Y, and G tr(Y) on the diagonal
    auto convert_elastic = [&]() __attribute__((always_inline)) {
#pragma unroll
      for (int u = 0; u < 2; u++)
#pragma unroll
        for (int v = 0; v < 2; v++) {
          double Y[3][3];
#pragma unroll
          for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) Y[i][j] = K[u][v][i][j];
          Dg[u][v] = Gm * (Y[0][0] + Y[1][1] + Y[2][2]);
#pragma unroll
          for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) K[u][v][i][j] = fma(Gm, Y[j][i], lamb * Y[i][j]);
        }
    };
#pragma unroll 1
    for (int c = 0; c < H_NCH; c++) {
      const int q0 = c * H_CH;
      lds_wait();  // coefficients (first chunk) written; the previous chunk's node vectors consumed
      if (stager) {
        const double* Q = sQ + (q0 + sql) * H_QW;
        const int q = (int)Q[18];
        const double* dN = sdN + (q * H_NN + sm) * H_ND;
        const double* Ji = Q + 9;
        double G[3];
#pragma unroll
        for (int j = 0; j < 3; j++) {
          double sum = 0.0;
#pragma unroll
          for (int k = 0; k < 3; k++) sum += dN[k] * Ji[k * 3 + j];
          G[j] = sum;
        }
        const double* sv = Q + 3;
        double* P = sP + (sql * H_NN + myrank) * H_PW;  // node m in the slot of its rank
        const double sxy = sv[3] * ISQ2_H, syz = sv[4] * ISQ2_H, sxz = sv[5] * ISQ2_H;
        P[0] = G[0]; P[1] = G[1]; P[2] = G[2];
        P[4] = sv[0] * G[0] + sxy * G[1] + sxz * G[2];
        P[5] = sxy * G[0] + sv[1] * G[1] + syz * G[2];
        P[6] = sxz * G[0] + syz * G[1] + sv[2] * G[2];
      }
      lds_wait();
      // the chunk's elastic slots first, then (after converting Y where the elastic slots end) its plastic ones
      const int nel_c = min(max(n_el - q0, 0), H_CH);  // uniform
      if (active) {
#pragma unroll 1
        for (int ql = 0; ql < nel_c; ql++) {  // elastic: Y += c G_a G_b^T
          const double cq = sQ[(q0 + ql) * H_QW];
          const double* Pq = sP + ql * H_NN * H_PW;
          double Ga[2][3];
#pragma unroll
          for (int u = 0; u < 2; u++)
#pragma unroll
            for (int i = 0; i < 3; i++) Ga[u][i] = Pq[(2 * I + u) * H_PW + i];
#pragma unroll
          for (int v = 0; v < 2; v++) {
            const double* Pb = Pq + (2 * J + v) * H_PW;
            double ub[3];
#pragma unroll
            for (int i = 0; i < 3; i++) ub[i] = cq * Pb[i];
#pragma unroll
            for (int u = 0; u < 2; u++)
#pragma unroll
              for (int i = 0; i < 3; i++)
#pragma unroll
                for (int j = 0; j < 3; j++) K[u][v][i][j] = fma(Ga[u][i], ub[j], K[u][v][i][j]);
          }
        }
        if (nel_c < H_CH && n_el >= q0) convert_elastic();  // the first plastic slot is in this chunk
#pragma unroll 1
        for (int ql = nel_c; ql < H_CH; ql++) {
          const double* Q = sQ + (q0 + ql) * H_QW;
          const double ha = Q[0], be = Q[1], ga = Q[2];
          const double* Pq = sP + ql * H_NN * H_PW;
          double Ga[2][3];
#pragma unroll
          for (int u = 0; u < 2; u++)
#pragma unroll
            for (int i = 0; i < 3; i++) Ga[u][i] = Pq[(2 * I + u) * H_PW + i];
#pragma unroll
          for (int v = 0; v < 2; v++) {
            const double* Pb = Pq + (2 * J + v) * H_PW;
            double ub[3], vb[3];
#pragma unroll
            for (int i = 0; i < 3; i++) {
              ub[i] = ha * Pb[i];
              vb[i] = be * Pb[i];
            }
#pragma unroll
            for (int u = 0; u < 2; u++)
#pragma unroll
              for (int i = 0; i < 3; i++) {
                Dg[u][v] = fma(Ga[u][i], ub[i], Dg[u][v]);
#pragma unroll
                for (int j = 0; j < 3; j++) {
                  double k = K[u][v][i][j];
                  k = fma(ub[i], Ga[u][j], k);
                  k = fma(Ga[u][i], vb[j], k);
                  K[u][v][i][j] = k;
                }
              }
          }
          // the s(x)s term only where the IP is plastic (gamma = 0 at elastic IPs): a branch uniform over the
          // wavefront (one element per wavefront) that only adds to the accumulators.  Every entry still takes
          // its alpha, beta and gamma terms of an IP in that order, as in the node-block kernel.
          if (ga != 0.0) {
            double ta[2][3];
#pragma unroll
            for (int u = 0; u < 2; u++)
#pragma unroll
              for (int i = 0; i < 3; i++) ta[u][i] = Pq[(2 * I + u) * H_PW + 4 + i];
#pragma unroll
            for (int v = 0; v < 2; v++) {
              const double* Pb = Pq + (2 * J + v) * H_PW;
              double wb[3];
#pragma unroll
              for (int i = 0; i < 3; i++) wb[i] = ga * Pb[4 + i];
#pragma unroll
              for (int u = 0; u < 2; u++)
#pragma unroll
                for (int i = 0; i < 3; i++)
#pragma unroll
                  for (int j = 0; j < 3; j++) K[u][v][i][j] = fma(ta[u][i], wb[j], K[u][v][i][j]);
            }
          }
        }
      }
      order();
    }
    if (n_el == H_NIP) convert_elastic();  // every IP elastic (uniform)
#pragma unroll
    for (int u = 0; u < 2; u++)
#pragma unroll
      for (int v = 0; v < 2; v++)
#pragma unroll
        for (int i = 0; i < 3; i++) K[u][v][i][i] += Dg[u][v];
    // output: the upper node blocks of this lane's tile at their stored offsets; a diagonal node block takes
    // its lower half from the upper one (every element matrix, and K, exactly symmetric)
    typedef KT V __attribute__((ext_vector_type(16 / sizeof(KT))));
    constexpr int VL = 16 / (int)sizeof(KT);
    constexpr int NPASS = sizeof(KT) == 8 ? 2 : 1;
    constexpr int PV = H_EW / NPASS;            // values per pass
    constexpr int NPIECE = PV / VL;             // 16-byte pieces per pass (473)
    constexpr int ROUNDS = (NPIECE + 63) / 64;  // 8
    KT* stage = reinterpret_cast<KT*>(sU);
    KT* dst0 = ke + e * H_EW;
#pragma unroll
    for (int p = 0; p < NPASS; p++) {
      lds_wait();  // the last chunk's node vectors (or the previous pass) consumed
      const bool mine = active && (NPASS == 1 || (I >= 3) == (p == 1));
      if (mine) {
#pragma unroll
        for (int u = 0; u < 2; u++)
#pragma unroll
          for (int v = 0; v < 2; v++) {
            const int a = 2 * I + u, b = 2 * J + v;
            if (a > b) continue;
            const int o = ke_block_offset(H_NN, H_ND, a, b) - p * PV;
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
              for (int j = 0; j < 3; j++) stage[o + i * 3 + j] = (KT)(a == b && j < i ? K[u][v][j][i] : K[u][v][i][j]);
          }
      }
      if (lane == 0) {  // padding values
        if (NPASS == 2) stage[H_HALF - 1] = (KT)0;
        else { stage[H_HALF - 1] = (KT)0; stage[H_EW - 1] = (KT)0; }
      }
      lds_wait();
      V* dst = reinterpret_cast<V*>(dst0 + p * PV);
      const V* src = reinterpret_cast<const V*>(stage);
#pragma unroll
      for (int r = 0; r < ROUNDS; r++) {
        const int i = min(lane + 64 * r, NPIECE - 1);  // surplus lanes repeat the last piece (same value)
        dst[i] = src[i];
      }
      order();
    }
    // the next element's data: the copies (and the ticket, issued first) precede this element's
    // NPASS * ROUNDS stores, and vector memory operations complete in order
    if constexpr (NPASS * ROUNDS == 16) asm volatile("s_waitcnt vmcnt(16)" ::: "memory");
    else asm volatile("s_waitcnt vmcnt(8)" ::: "memory");
    cur = nxt;
    asm volatile("" : "+v"(tk));  // read the ticket here, after the wait above
    nxt = cidx + (int64_t)ncnt * __builtin_amdgcn_readfirstlane(tk);
  }
}

__global__ void k_zero_work(int32_t* w) {
  if (threadIdx.x < kWorkCounters) w[threadIdx.x * kWorkStride] = 0;
}

}  // namespace

// LDS bytes of a persistent Hex20 tangent workgroup of nw wavefronts
static size_t hex20_lds(int nw) { return sizeof(double) * ((size_t)H_TAB + (size_t)nw * H_WAVE); }

// wavefronts per workgroup of the persistent Hex20 tangent on this device (0: it does not fit)
static int hex20_waves(const pm_fem* f) {
  int nw = H_WMAX;
  while (nw > 4 && (int64_t)hex20_lds(nw) > f->lds_per_block) nw--;
  return (int64_t)hex20_lds(nw) <= f->lds_per_block ? nw : 0;
}

// true when a group's vector tangent runs as the persistent Hex20 kernel: Hex20 cells with 27 IPs, FP64 or
// FP32 tangent, and a device whose workgroups may hold its LDS
bool tangent_hex20_fits(const pm_fem* f, const Group& g) {
  return g.kind == HEX20 && g.nip == H_NIP && !f->axisym && !g.plane_stress && f->cfg.tangent_precision != 2 &&
         hex20_waves(f) > 0;
}

// Launches the persistent Hex20 (27 IP) vector tangent of one group; false when the group or the device
// does not fit it (the caller then runs the node-block kernel).
bool launch_tangent_hex20(pm_fem* f, Group& g) {
  if (!tangent_hex20_fits(f, g)) return false;
  if (g.nel == 0) return true;
  const int nw = hex20_waves(f);
  const size_t shm = hex20_lds(nw);
  const int per_cu = std::max<int>(1, (int)(f->lds_per_cu / (int64_t)shm));
  const int64_t grid = std::min<int64_t>((int64_t)f->n_cu * per_cu, (g.nel + nw - 1) / nw);
  if (!ke_range_ok(g.ke_base, g.nel, g.kew, H_EW, f->n_ke_values)) {
    f->launch_refused = true;  // reported by check_device_error; never launched out of range
    return true;
  }
  H20Mat mp;
  mp.model = g.param.model; mp.K = g.K; mp.G = g.G; mp.H = g.param.hh;
  const uint64_t ekey = (uint64_t)(&g - f->groups.data()) << kErrKeyShift;
  // workgroups above 64 KB of LDS: allowed up to the device limit (the call is a formality where they already are)
  static thread_local bool attr_set = false;
  if (!attr_set) {
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_tangent_hex20<double>), hipFuncAttributeMaxDynamicSharedMemorySize, (int)f->lds_per_block);
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_tangent_hex20<float>), hipFuncAttributeMaxDynamicSharedMemorySize, (int)f->lds_per_block);
    (void)hipGetLastError();
    attr_set = true;
  }
  hipLaunchKernelGGL(k_zero_work, dim3(1), dim3(64), 0, f->stream, f->d_work);  // a kernel, not a memset (graph replays)
  if (f->cfg.tangent_precision == 0) {
    H20Args<double> A{g.nel, g.d_rank, g.d_geo, g.d_tab, g.d_state, f->d_ke + g.ke_base, mp, f->thickness, f->d_err, ekey, f->d_work,
                      (int)std::min<int64_t>(kWorkCounters, grid)};
    hipLaunchKernelGGL((k_tangent_hex20<double>), dim3((unsigned)grid), dim3(64 * nw), shm, f->stream, A);
  } else {
    H20Args<float> A{g.nel, g.d_rank, g.d_geo, g.d_tab, g.d_state, f->d_ke32 + g.ke_base, mp, f->thickness, f->d_err, ekey, f->d_work,
                     (int)std::min<int64_t>(kWorkCounters, grid)};
    hipLaunchKernelGGL((k_tangent_hex20<float>), dim3((unsigned)grid), dim3(64 * nw), shm, f->stream, A);
  }
  return true;
}

}  // namespace pm
