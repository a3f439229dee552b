// Persistent matrix-core tangent for Hex20 elements with 27 integration points: BF16 operands, FP32
// accumulation (the BF16 inexact-Newton tangent).
//
// For D = alpha I + beta I(x)I + gamma s(x)s (von_mises.rs:124-167) the element matrix is
//   K_ab = M_ab^T + tr(M_ab) 1 + C_ab,
//   M = sum_q (c alpha/2)_q g_q g_q^T,   C = sum_q (c beta)_q g_q g_q^T + (c gamma)_q ts_q ts_q^T,
// with g_q[(a,i)] = G_q[a][i] and ts_q = B_q^T s_q (the decomposition of the per-group matrix-core kernel).
// Each Gram is formed from a single row kind: the rows are scaled by the square roots of their weights
// (alpha > 0, beta >= 0 for nu >= 0), and the gamma term, whose weight is never positive, is subtracted
// by negating its B fragments in registers.  Node columns are padded to four (80 columns: 5x5 tiles of
// 16x16, 15 upper tiles), so that no 3x3 node block crosses a tile: the block transposes of M and its
// traces stay inside a tile, between the four lanes of a quad (DPP).  The IPs are taken in two chunks of 16
// (the MFMA depth), the rows of a chunk staged in LDS as BF16.
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
constexpr int M_KC = 16;                 // IPs per chunk (MFMA depth)
constexpr int M_NCHK = 2;                // chunks (IPs 27..31 are zero rows)
constexpr int M_RS = 20;                 // BF16 values per column row in LDS (16 used; 40 B: conflict-free reads)
constexpr int M_KIND = M_NC * M_RS;      // BF16 values per row kind
constexpr int M_QW = 9;                  // per IP: sqrt(c alpha/2), sqrt(c beta), sqrt(-c gamma), deviator (6)
constexpr int M_GW = 11;                 // per IP in the geometry array: J^-1 (9), det J, radius
constexpr int M_GEO_DW = 640, M_ST_DW = 448, M_RK_DW = 64;  // LDS copies of one element (dwords)
constexpr int M_RAW = (M_GEO_DW + M_ST_DW + M_RK_DW) / 2;    // doubles
constexpr int M_EW = packed_ke_values(M_NN, M_ND);           // 1892 values per element matrix
constexpr int M_ROWB = 3 * M_KIND * 2;                       // bytes of the three row kinds of a chunk
constexpr int M_STG = M_EW * 4;                              // output staging (FP32)
constexpr int M_UB = ((M_ROWB > M_STG ? M_ROWB : M_STG) + 15) / 16 * 16;  // union bytes
constexpr int M_WAVE = M_RAW + M_NIP * M_QW + 1 + M_UB / 8;  // doubles per wavefront (even)
constexpr int M_TAB = M_NIP * M_NN * M_ND + M_NIP + 1;       // dN, w, padding (doubles)
constexpr int M_WMAX = 8;
static_assert(M_WAVE % 2 == 0 && M_TAB % 2 == 0, "wavefront regions start 16-byte aligned");
static_assert(M_NIP <= M_KC * M_NCHK, "two chunks cover the IPs");

typedef __attribute__((address_space(3))) void lds_void_m;
typedef __attribute__((address_space(1))) int32_t gint_m;
typedef float f4_m __attribute__((ext_vector_type(4)));
typedef short s4_m __attribute__((ext_vector_type(4)));
typedef float V4 __attribute__((ext_vector_type(4)));
typedef float f2_m __attribute__((ext_vector_type(2)));
typedef __bf16 bf2_m __attribute__((ext_vector_type(2)));  // round to nearest even (v_cvt_pk_bf16_f32)

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

__global__ __launch_bounds__(64 * M_WMAX) void k_tangent_hex20_bf16(MArgs A) {
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
  const int nw = blockDim.x >> 6;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  double* sdN = sh;
  double* sw = sh + M_NIP * M_NN * M_ND;
  {
    const double* __restrict__ tab = A.tab;
    for (int i = tid; i < M_NIP * M_NN * M_ND; i += blockDim.x) sdN[i] = tab[M_NIP + M_NIP * M_NN + i];
    for (int i = tid; i < M_NIP; i += blockDim.x) sw[i] = tab[i];
  }
  __syncthreads();  // the only workgroup-wide wait
  const int ncnt = A.ncnt, cidx = (int)(blockIdx.x % (unsigned)ncnt);

  double* raw = sh + M_TAB + (size_t)wv * M_WAVE;
  double* rgeo = raw;
  double* rst = raw + M_GEO_DW / 2;
  double* rrk = rst + M_ST_DW / 2;
  const uint8_t* srk = reinterpret_cast<const uint8_t*>(rrk);
  double* sQ = raw + M_RAW;                                            // [27][9]
  char* uni = reinterpret_cast<char*>(sQ + M_NIP * M_QW + 1);          // rows of a chunk | output staging
  uint16_t* rows = reinterpret_cast<uint16_t*>(uni);                   // [3 kinds][80 columns][M_RS]
  float* stage = reinterpret_cast<float*>(uni);

  int32_t* work = A.work + cidx * kWorkStride;
  auto ticket = [&]() -> int32_t {
    int32_t t = 0;
    gint_m* wp = (gint_m*)work;
    asm volatile("" : "+v"(wp));
    if ((threadIdx.x & 63) == 0) t = __hip_atomic_fetch_add(wp, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return t;
  };
  auto prefetch_state = [&](int64_t e) {
    int lane = tid & 63;
    asm volatile("" : "+v"(lane));
#pragma unroll
    for (int k = 0; k < M_ST_DW / 64; k++) {
      const int d = min(64 * k + lane, 8 * M_NIP * 2 - 1);
      const int c = d / (2 * M_NIP), o = d - c * (2 * M_NIP);
      const int cg = c < M_NS ? c : c + 1;
      const uint32_t* src = reinterpret_cast<const uint32_t*>(st + cg * nipg + e * M_NIP) + o;
      __builtin_amdgcn_global_load_lds(src, (lds_void_m*)(rst + 32 * k), 4, 0, 0);
    }
    order_m();
  };
  auto prefetch_geo = [&](int64_t e, int k) {
    int lane = tid & 63;
    asm volatile("" : "+v"(lane));
    const uint32_t* src = reinterpret_cast<const uint32_t*>(geo + e * (M_NIP * M_GW));
    const int d = min(64 * k + lane, M_NIP * M_GW * 2 - 1);
    __builtin_amdgcn_global_load_lds(src + d, (lds_void_m*)(rgeo + 32 * k), 4, 0, 0);
    order_m();
  };
  auto prefetch_rank = [&](int64_t e) {
    int lane = tid & 63;
    asm volatile("" : "+v"(lane));
    const uint32_t* src = reinterpret_cast<const uint32_t*>(rank + e * M_NN);
    __builtin_amdgcn_global_load_lds(src + min(lane, M_NN / 4 - 1), (lds_void_m*)rrk, 4, 0, 0);
    order_m();
  };
  int32_t tk = ticket();
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  int64_t cur = cidx + (int64_t)ncnt * __builtin_amdgcn_readfirstlane(tk);
  if (cur >= nel) return;  // uniform over the wavefront
  prefetch_state(cur);
  for (int k = 0; k < M_GEO_DW / 64; k++) prefetch_geo(cur, k);
  prefetch_rank(cur);
  tk = ticket();
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  int64_t nxt = cidx + (int64_t)ncnt * __builtin_amdgcn_readfirstlane(tk);

  while (cur < nel) {
    const int64_t e = cur;
    int lane = tid & 63;
    asm volatile("" : "+v"(lane));
    // per IP: D coefficients (von_mises.rs:124-167) and deviator, one lane per IP; the Gram rows carry the
    // square roots of c alpha/2, c beta and -c gamma
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
      double* Q = sQ + q * M_QW;
      Q[0] = sqrt(coef * alpha * 0.5);
      Q[1] = sqrt(fmax(coef * beta, 0.0));
      Q[2] = sqrt(fmax(-coef * gamma, 0.0));
#pragma unroll
      for (int c = 0; c < 6; c++) Q[3 + c] = s[c];
    }
    int32_t tk2 = ticket();  // the element after nxt, read at the end of this element
    const bool next = nxt < nel;  // uniform
    if (next) {
      lds_wait_m();  // the state of this element has been read
      prefetch_state(nxt);
    }
    f4_m accM[15], accC[15];
#pragma unroll
    for (int t = 0; t < 15; t++) { accM[t] = f4_m{0, 0, 0, 0}; accC[t] = f4_m{0, 0, 0, 0}; }
#pragma unroll 1
    for (int h = 0; h < M_NCHK; h++) {
      lds_wait_m();  // coefficients written (first chunk); the previous chunk's rows consumed
      // rows of IPs 16h + 2p, 16h + 2p + 1 for node m (item p * 20 + m): three kinds x three dofs, the fourth
      // column of the node zero; IPs past 27 are zero rows
#pragma unroll 1
      for (int w = lane; w < (M_KC / 2) * M_NN; w += 64) {
        const int p = w / M_NN, m = w - p * M_NN;
        // the nine weighted row values (three kinds x three dofs) of each of the two IPs, in FP32; the products
        // of the weights with G and ts are formed in FP32 and rounded once to BF16 (two IPs per conversion)
        float v[2][9];
#pragma unroll
        for (int t = 0; t < 2; t++) {
          const int q = M_KC * h + 2 * p + t;
#pragma unroll
          for (int z = 0; z < 9; z++) v[t][z] = 0.0f;
          if (q < M_NIP) {
            const double* Q = sQ + q * M_QW;
            const double* dN = sdN + (q * M_NN + m) * M_ND;
            const double* Ji = rgeo + q * M_GW;
            double G[3];
#pragma unroll
            for (int j = 0; j < 3; j++) {
              double sum = 0.0;
#pragma unroll
              for (int k = 0; k < 3; k++) sum += dN[k] * Ji[k * 3 + j];
              G[j] = sum;
            }
            const double* sv = Q + 3;
            const double sxy = sv[3] * ISQ2_M, syz = sv[4] * ISQ2_M, sxz = sv[5] * ISQ2_M;
            const float ts0 = (float)(sv[0] * G[0] + sxy * G[1] + sxz * G[2]);
            const float ts1 = (float)(sxy * G[0] + sv[1] * G[1] + syz * G[2]);
            const float ts2 = (float)(sxz * G[0] + syz * G[1] + sv[2] * G[2]);
            const float wa = (float)Q[0], wb = (float)Q[1], wg = (float)Q[2];
#pragma unroll
            for (int i = 0; i < 3; i++) {
              const float gi = (float)G[i];
              v[t][i] = wa * gi;
              v[t][3 + i] = wb * gi;
            }
            v[t][6] = wg * ts0; v[t][7] = wg * ts1; v[t][8] = wg * ts2;
          }
        }
        const int col = 4 * srk[m];
#pragma unroll
        for (int z = 0; z < 3; z++) {
          uint32_t* r32 = reinterpret_cast<uint32_t*>(rows + z * M_KIND + col * M_RS + 2 * p);
#pragma unroll
          for (int i = 0; i < 3; i++) {
            const f2_m pr = {v[0][3 * z + i], v[1][3 * z + i]};
            r32[i * (M_RS / 2)] = __builtin_bit_cast(uint32_t, __builtin_convertvector(pr, bf2_m));
          }
          r32[3 * (M_RS / 2)] = 0u;
        }
      }
      lds_wait_m();
      if (next) {  // the J^-1 of this chunk's IPs has been read (and, after the last chunk, the ranks)
        if (h == 0) { for (int k = 0; k < 5; k++) prefetch_geo(nxt, k); }
        else { for (int k = 5; k < M_GEO_DW / 64; k++) prefetch_geo(nxt, k); prefetch_rank(nxt); }
      }
      // Gram products of the chunk on the upper tiles: A fragment = 16 columns of tile I x 4 IPs of this
      // lane's group, B fragment likewise for tile J (negated for the gamma term)
      const int fo = (lane & 15) * M_RS + 4 * (lane >> 4);
      int t = 0;
#pragma unroll
      for (int I = 0; I < M_T; I++) {
        const s4_m a1 = *reinterpret_cast<const s4_m*>(rows + 0 * M_KIND + 16 * I * M_RS + fo);
        const s4_m a2 = *reinterpret_cast<const s4_m*>(rows + 1 * M_KIND + 16 * I * M_RS + fo);
        const s4_m a3 = *reinterpret_cast<const s4_m*>(rows + 2 * M_KIND + 16 * I * M_RS + fo);
#pragma unroll
        for (int J = I; J < M_T; J++, t++) {
          const s4_m b1 = *reinterpret_cast<const s4_m*>(rows + 0 * M_KIND + 16 * J * M_RS + fo);
          const s4_m b2 = *reinterpret_cast<const s4_m*>(rows + 1 * M_KIND + 16 * J * M_RS + fo);
          s4_m b3 = *reinterpret_cast<const s4_m*>(rows + 2 * M_KIND + 16 * J * M_RS + fo);
          b3 ^= (short)0x8000;  // -sqrt(-c gamma) ts: subtracts the gamma Gram
          accM[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a1, b1, accM[t], 0, 0, 0);
          accC[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a2, b2, accC[t], 0, 0, 0);
          accC[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a3, b3, accC[t], 0, 0, 0);
        }
      }
      order_m();
    }
    // epilogue per tile: K = M^T (per node block) + tr(M) 1 + C into the staging at the stored offsets of the
    // upper node blocks.  Lane: column j = 4 bp + p of tile J, rows 4g..4g+3 of tile I, so the four lanes of a
    // quad hold one 4x4 node block of M by columns; its transpose and trace move between them by DPP (the
    // padding row and column of a node are zero, so the trace over four is the trace over three).
    lds_wait_m();  // the rows are consumed (the staging overlays them)
    {
      const int j = lane & 15, g = lane >> 4, bp = j >> 2, p = j & 3;
      int t = 0;
#pragma unroll
      for (int I = 0; I < M_T; I++)
#pragma unroll
        for (int J = I; J < M_T; J++, t++) {
          const float x[4] = {accM[t][0], accM[t][1], accM[t][2], accM[t][3]};
          float T[4];  // T[s] = M(row 4g + p, col 4bp + (p ^ s)): entry (p, p ^ s) of the block, transposed
          T[0] = pick4(x, p);
          T[1] = quad_xor<1>(pick4(x, p ^ 1));
          T[2] = quad_xor<2>(pick4(x, p ^ 2));
          T[3] = quad_xor<3>(pick4(x, p ^ 3));
          const float t1 = T[0] + quad_xor<1>(T[0]);
          const float tr = t1 + quad_xor<2>(t1);
          const int a = 4 * I + g, b = 4 * J + bp;
          if (p < 3 && (I < J || g <= bp)) {
            const int o = ke_block_offset(M_NN, M_ND, a, b);
#pragma unroll
            for (int i = 0; i < 3; i++) {
              const float v = accC[t][i] + pick4(T, i ^ p);
              stage[o + i * 3 + p] = i == p ? v + tr : v;
            }
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

static size_t bf16_lds(int nw) { return sizeof(double) * ((size_t)M_TAB + (size_t)nw * M_WAVE); }

static int bf16_waves(const pm_fem* f) {
  int nw = M_WMAX;
  while (nw > 2 && (int64_t)bf16_lds(nw) > f->lds_per_block) nw--;
  return (int64_t)bf16_lds(nw) <= f->lds_per_block ? nw : 0;
}

}  // namespace

// true when a group's BF16 tangent runs as the persistent Hex20 matrix-core kernel: Hex20 cells with 27 IPs,
// a Poisson ratio >= 0 (beta >= 0, so that every Gram weight has a real square root), and a device whose
// workgroups may hold its LDS
bool tangent_hex20_bf16_fits(const pm_fem* f, const Group& g) {
  return g.kind == HEX20 && g.nip == M_NIP && !f->axisym && !g.plane_stress && f->cfg.tangent_precision == 2 &&
         g.param.poisson >= 0.0 && bf16_waves(f) > 0;
}

bool launch_tangent_hex20_bf16(pm_fem* f, Group& g) {
  if (!tangent_hex20_bf16_fits(f, g)) return false;
  if (g.nel == 0) return true;
  const int nw = bf16_waves(f);
  const size_t shm = bf16_lds(nw);
  const int per_cu = std::max<int>(1, (int)(f->lds_per_cu / (int64_t)shm));
  const int64_t grid = std::min<int64_t>((int64_t)f->n_cu * per_cu, (g.nel + nw - 1) / nw);
  if (!ke_range_ok(g.ke_base, g.nel, g.kew, M_EW, f->n_ke_values)) {
    f->launch_refused = true;  // reported by check_device_error; never launched out of range
    return true;
  }
  static thread_local bool attr_set = false;
  if (!attr_set) {
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_tangent_hex20_bf16), hipFuncAttributeMaxDynamicSharedMemorySize, (int)f->lds_per_block);
    (void)hipGetLastError();
    attr_set = true;
  }
  MMat mp;
  mp.model = g.param.model; mp.K = g.K; mp.G = g.G; mp.H = g.param.hh;
  const uint64_t ekey = (uint64_t)(&g - f->groups.data()) << kErrKeyShift;
  hipLaunchKernelGGL(k_zero_work_m, dim3(1), dim3(64), 0, f->stream, f->d_work);  // a kernel, not a memset (graph replays)
  MArgs A{g.nel, g.d_rank, g.d_geo, g.d_tab, g.d_state, f->d_ke32 + g.ke_base, mp, f->thickness, f->d_err, ekey, f->d_work,
          (int)std::min<int64_t>(kWorkCounters, grid)};
  hipLaunchKernelGGL(k_tangent_hex20_bf16, dim3((unsigned)grid), dim3(64 * nw), shm, f->stream, A);
  return true;
}

}  // namespace pm
