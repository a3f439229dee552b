// Element tangent on the matrix cores (MFMA), FP64 / FP32 / BF16-with-FP32-accumulation.
//
// For D = alpha*I + beta*I(x)I + gamma*s(x)s (Mandel; von_mises.rs:124-167) the element matrix is
//   K_ab = M_ab^T + tr(M_ab) 1 + C_ab,
//   M = sum_q (c alpha/2)_q  g_q g_q^T                (g_q[(a,i)] = G_q[a][i], the gradient vector)
//   C = sum_q (c beta)_q tI_q tI_q^T + (c gamma)_q ts_q ts_q^T (+ (c alpha)_q h_q h_q^T, axisymmetric)
// with tI = B^T I (= g unless axisymmetric), ts = B^T s and h[(a,0)] = N_a / r.  Both are Gram products
// over rows built per IP, i.e. small GEMMs (NN*ND x NN*ND, k = IPs) that run as 16x16 MFMA tiles; only
// the upper tiles are computed.  The rows are laid out with the nodes in rank order (sorted by global
// id), so the product comes out directly in the order the deterministic scatter reads.
//
// Organisation (persistent): each workgroup owns a contiguous range of units of EPW elements and keeps
// the shape-function table in LDS.  While one unit is processed, the IP geometry (J^-1, det J, radius,
// formed once for the fixed mesh), state and ranks of the next unit are already in flight in registers.
// Per element: one lane per IP forms the D coefficients; all threads of the element stage the gradient
// rows; WPE waves share the upper tiles (tile t on wave t % WPE); the epilogue parks M in LDS, forms K,
// stages the node blocks in LDS and writes the element matrix contiguously.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "pm_internal.hpp"

namespace pm {

namespace {

constexpr double ISQ2 = 0.70710678118654752440;

// material constants, with the products the D coefficients need formed on the host (they then stay in
// scalar registers instead of occupying vector registers across the persistent loop)
struct MatM {
  int model, plane_stress;
  double G3;      // 3 G
  double G2;      // 2 G
  double G6G;     // 6 G^2
  double inv_d;   // 1 / (3 G + H)
  double K;       // bulk modulus
  double beta_e;  // K - 2G/3 (elastic)
  double a_ps, b_ps;  // plane stress: E/(1+nu), E nu/(1-nu^2)
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
template <> struct MfmaT<0> {  // FP64 16x16x4: output register r of lane l holds D[4r + l/16][l%16]
  typedef double T; typedef d4_t Acc;
  static constexpr int KS = 4;
  __device__ static int row(int l, int r) { return 4 * r + (l >> 4); }
};
template <> struct MfmaT<1> {  // FP32 16x16x4: output register r of lane l holds D[4(l/16) + r][l%16]
  typedef float T; typedef f4_t Acc;
  static constexpr int KS = 4;
  __device__ static int row(int l, int r) { return 4 * (l >> 4) + r; }
};
template <> struct MfmaT<2> {  // BF16 16x16x16 with FP32 accumulation, same output layout as FP32
  typedef float T; typedef f4_t Acc;
  static constexpr int KS = 16;
  __device__ static int row(int l, int r) { return 4 * (l >> 4) + r; }
};

template <int ND, int NN>
struct MShape {
  static constexpr int NS = ND == 2 ? 4 : 6;
  static constexpr int NC = NN * ND;
  static constexpr int NTI = (NC + 15) / 16;
  static constexpr int NCP = NTI * 16;
  static constexpr int NTILE = NTI * (NTI + 1) / 2;
  static constexpr int WPE = NTILE == 1 ? 1 : (NTILE <= 4 ? NTILE : 4);  // waves per element
  static constexpr int EPW = WPE == 1 ? 4 : 1;                           // elements per workgroup (a unit)
  static constexpr int TPW = (NTILE + WPE - 1) / WPE;                     // tiles per wave
  static constexpr int TPE = 64 * WPE;                                    // threads per element
  static constexpr int NTHR = TPE * EPW;
  static constexpr int QW = 3 + ND * ND + 6 + 1;  // c alpha, c beta, c gamma, J^-1, s, r
  static constexpr int SV = NS + 2;               // state values read per IP: sigma, lambda, elastic flag (z skipped)
  static constexpr int GW = ND * ND + 2;          // geometry values per IP: J^-1, det J, radius
};

// waves per SIMD the register budget is sized for: four (128 VGPRs) except for the 3D FP64 kernels
template <int ND, int PREC> constexpr int mfma_minw() { return ND == 3 && PREC == 0 ? 3 : 4; }

// LDS, in doubles: [table: w | N (axisymmetric)] then EPW element slots of
// [state | per-IP data | union(IP geometry until the per-IP data is formed, rows + weights of a chunk,
// M tiles, node-block staging)]
struct MLayout {
  int rsk, nrow, dsn;            // row-slot stride per column (one row kind), row kinds, dN stride per IP
  size_t tdn, tn, tab;           // table offsets and size
  size_t s, q, uni, per_el;      // element-slot offsets and size
  size_t total;
};

template <int ND, int NN, int PREC>
__host__ __device__ inline MLayout mfma_layout(int nip, int ipc, int axisym) {
  using S = MShape<ND, NN>;
  using T = typename MfmaT<PREC>::T;
  MLayout L;
  if (PREC == 2) {
    // BF16 operands converted once when the rows are staged: g, ts and the weighted g wM, g wB, ts wG
    // (axisymmetric: g, tI, ts, h and g wM, tI wB, ts wG, h wH); 4-value groups 8-byte aligned
    L.nrow = axisym ? 8 : 5;
    L.rsk = ipc + 4;
  } else {
    L.nrow = axisym ? 4 : 2;  // g, ts  |  g, tI, ts, h
    L.rsk = ipc + 1;          // ipc is a multiple of 4: odd stride
  }
  L.dsn = NN * ND + 1;
  L.tdn = (size_t)nip;
  L.tn = L.tdn;  // dN is read through the L1 (it is shared by every workgroup)
  L.tab = L.tn + (axisym ? (size_t)nip * NN : 0);
  L.tab = (L.tab + 1) & ~(size_t)1;  // element slots and their union start 16-byte aligned
  L.s = 0;
  L.q = L.s + (size_t)nip * S::SV;
  L.uni = (L.q + (size_t)nip * S::QW + 1) & ~(size_t)1;
  const size_t rows = PREC == 2 ? (size_t)L.nrow * S::NCP * L.rsk * sizeof(short)
                                 : ((size_t)L.nrow * S::NCP * L.rsk + 4 * (size_t)ipc) * sizeof(T);
  const size_t mt = (size_t)S::NCP * (S::NCP + 1) * sizeof(T);  // dense M, both triangles
  const size_t kst = (size_t)packed_ke_values(NN, ND) * sizeof(T);
  size_t u = rows > mt ? rows : mt;
  u = u > kst ? u : kst;
  u = u > (size_t)nip * S::GW * sizeof(double) ? u : (size_t)nip * S::GW * sizeof(double);
  L.per_el = (L.uni + (u + sizeof(double) - 1) / sizeof(double) + 1) & ~(size_t)1;
  L.total = L.tab + (size_t)S::EPW * L.per_el;
  return L;
}

template <int ND, int NN, int PREC>
__global__ __launch_bounds__((MShape<ND, NN>::NTHR), (mfma_minw<ND, PREC>())) void k_tangent_mfma(
    int64_t nel, int nip, int ipc, int64_t units_per_wg, const uint8_t* __restrict__ rank,
    const double* __restrict__ geo, const double* __restrict__ tab, const double* __restrict__ st,
    typename MfmaT<PREC>::T* __restrict__ ke, MatM mp, int axisym, double thickness, int32_t* err, uint64_t ekey) {
  using S = MShape<ND, NN>;
  using MT = MfmaT<PREC>;
  using T = typename MT::T;
  using Acc = typename MT::Acc;
  constexpr int NS = S::NS, NC = S::NC, NTI = S::NTI, NCP = S::NCP, NTILE = S::NTILE, WPE = S::WPE, EPW = S::EPW;
  constexpr int TPW = S::TPW, TPE = S::TPE, NTHR = S::NTHR, QW = S::QW, SV = S::SV, GW = S::GW;
  constexpr int KS = MT::KS, B2 = ND * ND, EW = packed_ke_values(NN, ND), EU = NN * (NN + 1) / 2 * B2;
  extern __shared__ double shm[];
  __shared__ uint8_t srank[EPW * NN];
  const MLayout L = mfma_layout<ND, NN, PREC>(nip, ipc, axisym);
  const int RSK = L.rsk, NR = L.nrow;
  const int tid = threadIdx.x, wave = tid >> 6;
  (void)wave;
  const int64_t nunits = (nel + EPW - 1) / EPW;
  const int64_t u0 = (int64_t)blockIdx.x * units_per_wg;
  const int64_t u1 = min(nunits, u0 + units_per_wg);
  if (u0 >= u1) return;  // uniform over the workgroup
  const int64_t nipg = nel * nip;

  double* sw = shm;
  const double* __restrict__ tdN = tab + nip + nip * NN;
  double* sN = shm + L.tn;
  for (int i = tid; i < nip; i += NTHR) sw[i] = tab[i];
  if (axisym) for (int i = tid; i < nip * NN; i += NTHR) sN[i] = tab[nip + i];

  // prefetch roles (fixed per thread): up to two geometry values (element ge0 / ge1, both units of the
  // unit's contiguous geometry; EPW * nip * GW <= 2 NTHR, checked on the host), one rank (thread < EPW NN)
  // and one state value (se, component, q)
  const int gpe = nip * GW;
  const int ge0 = tid / gpe, ge1 = (tid + NTHR) / gpe;
  const bool grole0 = ge0 < EPW, grole1 = ge1 < EPW;
  const int gl0 = ge0 * (int)L.per_el + (int)L.uni + (tid - ge0 * gpe);
  const int gl1 = ge1 * (int)L.per_el + (int)L.uni + (tid + NTHR - ge1 * gpe);
  const bool rrole = tid < EPW * NN;
  const int re = tid / NN;
  const bool srole = tid < EPW * SV * nip;
  const int se = tid / (SV * nip), sc = (tid - se * SV * nip) / nip;
  const int64_t soff = (int64_t)(sc < NS ? sc : sc + 1) * nipg + (tid - se * SV * nip - sc * nip) + (int64_t)se * nip;  // skips z
  const int sl = se * (int)(L.per_el) + (int)L.s + (tid - se * SV * nip);  // LDS target of the state value
  double rg0 = 0.0, rg1 = 0.0, rs = 0.0;
  uint8_t rr = 0;
  // The prefetch loads are issued unconditionally (surplus lanes re-read a valid entry) so the number of
  // memory operations after each is static and waiting for a prefetch never waits for earlier stores.
  // Offsets pass through an empty asm so that no 64-bit address is kept live, and spilled, across units.
  auto load_unit = [&](int64_t u) {
    const int n = (int)min((int64_t)EPW, nel - u * EPW);
    int go0 = (grole0 && ge0 < n) ? tid : 0, go1 = (grole1 && ge1 < n) ? tid + NTHR : 0;
    int ro = (rrole && re < n) ? tid : 0;
    int64_t so = (srole && se < n) ? soff : 0;
    asm volatile("" : "+v"(go0), "+v"(go1), "+v"(ro), "+v"(so));
    const double* gu = geo + u * (EPW * (int64_t)gpe);
    rg0 = gu[go0];
    rg1 = gu[go1];
    rr = rank[u * (EPW * NN) + ro];
    rs = st[so + u * (EPW * (int64_t)nip)];
  };
  load_unit(u0);

  for (int64_t u = u0; u < u1; u++) {
    const int nloc = (int)min((int64_t)EPW, nel - u * EPW);
    __syncthreads();  // previous unit finished with LDS (and the table is in place)
    if (grole0 && ge0 < nloc) shm[L.tab + gl0] = rg0;
    if (grole1 && ge1 < nloc) shm[L.tab + gl1] = rg1;
    if (rrole && re < nloc) srank[tid] = rr;
    if (srole && se < nloc) shm[L.tab + sl] = rs;
    load_unit(min(u + 1, u1 - 1));  // past the range: a harmless repeat
    __syncthreads();
    // Per-thread indices are re-derived each unit from an opaque copy of the thread id, so the compiler
    // does not hoist every loop-invariant LDS address out of the unit loop (that would exhaust registers).
    int tq = tid;
    asm volatile("" : "+v"(tq));
    const int l = tq & 63, el = (tq >> 6) / WPE, wv = (tq >> 6) - el * WPE, te = tq - el * TPE;
    const bool active = EPW == 1 || el < nloc;  // a unit of one element is always full
    double* slot = shm + L.tab + el * L.per_el;  // this thread's element slot
    const double* sG = slot + L.uni;  // IP geometry: [nip][GW], until the per-IP data is formed
    double* sS = slot + L.s;  // state: [SV][nip]
    double* sQ = slot + L.q;
    T* sL = reinterpret_cast<T*>(slot + L.uni);  // row kind k: sL[(k * NCP + col) * RSK + slot]; weights after
    int tIv[TPW], tJv[TPW];
#pragma unroll
    for (int s = 0; s < TPW; s++) {
      int t = wv + WPE * s, I = 0;
      if (t >= NTILE) t = NTILE - 1;  // idle slot: computed, never stored
      while (t >= NTI - I) { t -= NTI - I; I++; }
      tIv[s] = I; tJv[s] = I + t;
    }

    // one lane per IP: c times the D coefficients (von_mises.rs:124-167), deviator; J^-1 and radius copied
    if (active && wv == 0 && l < nip) {
      const int q = l;
      const double* gq = sG + q * GW;
      const double det = gq[ND * ND], r = gq[ND * ND + 1];
      const uint64_t ikey = ekey + (uint64_t)((u * EPW + el) * nip + q);
      if (!(det > 0.0)) raise_error(err, ERR_DETJ, ikey);
      const double coef = det * sw[q] * thickness * (axisym ? r : 1.0);
      double alpha, beta, gamma = 0.0;
      double s[6] = {0, 0, 0, 0, 0, 0};
      if (mp.plane_stress) {
        alpha = mp.a_ps;
        beta = mp.b_ps;
      } else if (mp.model == 0 || sS[(NS + 1) * nip + q] != 0.0) {
        alpha = mp.G2;
        beta = mp.beta_e;
      } else {
        double sig[NS];
#pragma unroll
        for (int c = 0; c < NS; c++) sig[c] = sS[c * nip + q];
        const double lam = sS[NS * nip + q];
        const double sm = (sig[0] + sig[1] + sig[2]) / 3.0;
        double ss = 0.0;
#pragma unroll
        for (int c = 0; c < NS; c++) { s[c] = sig[c] - (c < 3 ? sm : 0.0); ss += s[c] * s[c]; }
        const double sd = sqrt(1.5 * ss);
        const double sdtr = sd + lam * mp.G3;
        const double norm_s = sd * 0.81649658092772603273;  // sqrt(2/3)
        if (norm_s < 1e-10) raise_error(err, ERR_NORMS, ikey);
        alpha = mp.G2 * (1.0 - lam * mp.G3 / sdtr);
        gamma = mp.G6G * (lam / sdtr - mp.inv_d) / (norm_s * norm_s);
        beta = mp.K - alpha / 3.0;
      }
      double* Q = sQ + q * QW;
      Q[0] = coef * alpha;
      Q[1] = coef * beta;
      Q[2] = coef * gamma;
#pragma unroll
      for (int ij = 0; ij < ND * ND; ij++) Q[3 + ij] = gq[ij];
#pragma unroll
      for (int c = 0; c < 6; c++) Q[3 + ND * ND + c] = s[c];
      Q[3 + ND * ND + 6] = r;
    }

    Acc accM[TPW], accC[TPW];
#pragma unroll
    for (int s = 0; s < TPW; s++) { accM[s] = Acc{0, 0, 0, 0}; accC[s] = Acc{0, 0, 0, 0}; }
    const int KR = NCP * RSK;  // values per row kind
    T* sW = sL + NR * KR;      // weights: 4 kinds x ipc
    const int nchunk = (nip + ipc - 1) / ipc;
    for (int ch = 0; ch < nchunk; ch++) {
      __syncthreads();  // per-IP data ready / previous chunk consumed
      const int nv = min(ipc, nip - ch * ipc);  // valid IPs of this chunk
      if (active && PREC == 2) {
        // BF16 rows of the valid IPs (nodes in rank order, column ND * rank(m) + i), each operand
        // rounded once here: the unweighted A rows and the B rows times their IP weight (the weight
        // product in FP32, then rounded, as the MFMA operand)
        short* sB = reinterpret_cast<short*>(slot + L.uni);
        const int KB = NCP * RSK;
        for (int w = te; w < nv * NN; w += TPE) {
          const int ql = w / NN, m = w - ql * NN, q = ch * ipc + ql;
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
          const double h = axisym ? sN[q * NN + m] / Q[3 + ND * ND + 6] : 0.0;
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
          const float wM = (float)(0.5 * Q[0]), wB = (float)Q[1], wG = (float)Q[2], wH = (float)Q[0];
          short* c = sB + (ND * srank[el * NN + m]) * RSK + ql;
#pragma unroll
          for (int i = 0; i < ND; i++) {
            const float g = (float)G[i], t = (float)ts[i];
            if (NR == 5) {
              c[i * RSK] = to_bf16(g);
              c[KB + i * RSK] = to_bf16(t);
              c[2 * KB + i * RSK] = to_bf16(g * wM);
              c[3 * KB + i * RSK] = to_bf16(g * wB);
              c[4 * KB + i * RSK] = to_bf16(t * wG);
            } else {
              const float ti = (float)(G[i] + (i == 0 ? h : 0.0)), hh = i == 0 ? (float)h : 0.0f;
              c[i * RSK] = to_bf16(g);
              c[KB + i * RSK] = to_bf16(ti);
              c[2 * KB + i * RSK] = to_bf16(t);
              c[3 * KB + i * RSK] = to_bf16(hh);
              c[4 * KB + i * RSK] = to_bf16(g * wM);
              c[5 * KB + i * RSK] = to_bf16(ti * wB);
              c[6 * KB + i * RSK] = to_bf16(t * wG);
              c[7 * KB + i * RSK] = to_bf16(hh * wH);
            }
          }
        }
        // zero padding: slots of IPs beyond nip (all columns), and columns [NC, NCP) of the valid slots
        const int npad_q = (ipc - nv) * NCP, npad = npad_q + (NCP - NC) * nv;
        for (int w = te; w < npad; w += TPE) {
          int col, ql;
          constexpr int PADC = NCP > NC ? NCP - NC : 1;
          if (w < npad_q) { ql = nv + w / NCP; col = w - (ql - nv) * NCP; }
          else { const int w2 = w - npad_q; ql = w2 / PADC; col = NC + (w2 - ql * PADC); }
          for (int k = 0; k < NR; k++) sB[k * KB + col * RSK + ql] = 0;
        }
      } else if (active) {
        // rows of the valid IPs, nodes in rank order: column ND * rank(m) + i
        for (int w = te; w < nv * NN; w += TPE) {
          const int ql = w / NN, m = w - ql * NN, q = ch * ipc + ql;
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
          const double h = axisym ? sN[q * NN + m] / Q[3 + ND * ND + 6] : 0.0;
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
          T* c = sL + (ND * srank[el * NN + m]) * RSK + ql;
#pragma unroll
          for (int i = 0; i < ND; i++) {
            c[i * RSK] = (T)G[i];
            if (NR == 2) {
              c[KR + i * RSK] = (T)ts[i];
            } else {
              c[KR + i * RSK] = (T)(G[i] + (i == 0 ? h : 0.0));
              c[2 * KR + i * RSK] = (T)ts[i];
              c[3 * KR + i * RSK] = i == 0 ? (T)h : T(0);
            }
          }
        }
        // zero padding: slots of IPs beyond nip (all columns), and columns [NC, NCP) of the valid slots
        const int npad_q = (ipc - nv) * NCP, npad = npad_q + (NCP - NC) * nv;
        for (int w = te; w < npad; w += TPE) {
          int col, ql;
          constexpr int PADC = NCP > NC ? NCP - NC : 1;
          if (w < npad_q) { ql = nv + w / NCP; col = w - (ql - nv) * NCP; }
          else { const int w2 = w - npad_q; ql = w2 / PADC; col = NC + (w2 - ql * PADC); }
          for (int k = 0; k < NR; k++) sL[k * KR + col * RSK + ql] = T(0);
        }
        for (int ql = te; ql < ipc; ql += TPE) {
          const bool ok = ql < nv;
          const double* Q = sQ + (ok ? ch * ipc + ql : 0) * QW;
          sW[ql] = ok ? (T)(0.5 * Q[0]) : T(0);        // g rows into M: c alpha / 2
          sW[ipc + ql] = ok ? (T)Q[1] : T(0);          // tI rows: c beta
          sW[2 * ipc + ql] = ok ? (T)Q[2] : T(0);      // ts rows: c gamma
          sW[3 * ipc + ql] = ok ? (T)Q[0] : T(0);      // h rows: c alpha
        }
      }
      __syncthreads();
      if (active) {
        const int col = l & 15;
        const int tIk = NR == 2 ? 0 : KR;        // the tI rows (g itself unless axisymmetric)
        const int tsk = NR == 2 ? KR : 2 * KR;   // the ts rows
        if constexpr (PREC == 2) {
          // operands as staged: one 8-byte read per operand and tile (kinds: A rows first, then B rows)
          const short* sB = reinterpret_cast<const short*>(slot + L.uni);
          const int KB = NCP * RSK;
          const int kA_i = NR == 5 ? 0 : 1, kA_t = NR == 5 ? 1 : 2, kB = NR == 5 ? 2 : 4;
#pragma unroll 1
          for (int k0 = 0; k0 < ipc; k0 += KS) {
            const int kq = k0 + 4 * (l >> 4);
#pragma unroll
            for (int s = 0; s < TPW; s++) {
              const short* ca = sB + (16 * tIv[s] + col) * RSK + kq;
              const short* cb = sB + (16 * tJv[s] + col) * RSK + kq;
              const s4_t ga = *reinterpret_cast<const s4_t*>(ca);
              const s4_t gm = *reinterpret_cast<const s4_t*>(cb + kB * KB);
              const s4_t ia = *reinterpret_cast<const s4_t*>(ca + kA_i * KB);
              const s4_t gb = *reinterpret_cast<const s4_t*>(cb + (kB + 1) * KB);
              const s4_t ta = *reinterpret_cast<const s4_t*>(ca + kA_t * KB);
              const s4_t tb = *reinterpret_cast<const s4_t*>(cb + (kB + 2) * KB);
              accM[s] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ga, gm, accM[s], 0, 0, 0);
              accC[s] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ia, gb, accC[s], 0, 0, 0);
              accC[s] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ta, tb, accC[s], 0, 0, 0);
              if (NR == 8) {
                const s4_t ha = *reinterpret_cast<const s4_t*>(ca + 3 * KB);
                const s4_t hb = *reinterpret_cast<const s4_t*>(cb + 7 * KB);
                accC[s] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ha, hb, accC[s], 0, 0, 0);
              }
            }
          }
        } else {
#pragma unroll 1
          for (int k0 = 0; k0 < ipc; k0 += KS) {
            const int k = k0 + (l >> 4);
            const T wm = sW[k], wb = sW[ipc + k], wg = sW[2 * ipc + k];
#pragma unroll
            for (int s = 0; s < TPW; s++) {
              const T* ca = sL + (16 * tIv[s] + col) * RSK;
              const T* cb = sL + (16 * tJv[s] + col) * RSK;
              const T ga = ca[k], gbv = cb[k], ia = ca[tIk + k], ib = cb[tIk + k] * wb, ta = ca[tsk + k], tb = cb[tsk + k] * wg;
              if constexpr (PREC == 0) {
                accM[s] = __builtin_amdgcn_mfma_f64_16x16x4f64(ga, gbv * wm, accM[s], 0, 0, 0);
                accC[s] = __builtin_amdgcn_mfma_f64_16x16x4f64(ia, ib, accC[s], 0, 0, 0);
                accC[s] = __builtin_amdgcn_mfma_f64_16x16x4f64(ta, tb, accC[s], 0, 0, 0);
                if (NR == 4) accC[s] = __builtin_amdgcn_mfma_f64_16x16x4f64(ca[3 * KR + k], cb[3 * KR + k] * sW[3 * ipc + k], accC[s], 0, 0, 0);
              } else {
                accM[s] = __builtin_amdgcn_mfma_f32_16x16x4f32(ga, gbv * wm, accM[s], 0, 0, 0);
                accC[s] = __builtin_amdgcn_mfma_f32_16x16x4f32(ia, ib, accC[s], 0, 0, 0);
                accC[s] = __builtin_amdgcn_mfma_f32_16x16x4f32(ta, tb, accC[s], 0, 0, 0);
                if (NR == 4) accC[s] = __builtin_amdgcn_mfma_f32_16x16x4f32(ca[3 * KR + k], cb[3 * KR + k] * sW[3 * ipc + k], accC[s], 0, 0, 0);
              }
            }
          }
        }
      }
    }
    __syncthreads();  // rows consumed
    // epilogue 1: park M densely (both triangles, row stride MDS)
    constexpr int MDS = NCP + 1;
    T* sM = sL;
    if (active) {
#pragma unroll
      for (int s = 0; s < TPW; s++) {
        if (wv + WPE * s >= NTILE) continue;
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int R = 16 * tIv[s] + MT::row(l, r), Cc = 16 * tJv[s] + (l & 15);
          sM[R * MDS + Cc] = accM[s][r];
          sM[Cc * MDS + R] = accM[s][r];
        }
      }
    }
    __syncthreads();
    // epilogue 2: K = M^T (per node block) + tr(M_ab) 1 + C, kept in the C accumulators.  For entry
    // (R, Cc) = (ND a + i, ND b + j): M^T is M[R - i + j][Cc - j + i], the trace runs down the block diagonal.
    if (active) {
#pragma unroll
      for (int s = 0; s < TPW; s++) {
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int R = 16 * tIv[s] + MT::row(l, r), Cc = 16 * tJv[s] + (l & 15);
          if (R >= NC || Cc >= NC) continue;
          const int i = R % ND, j = Cc % ND;
          const int base = R * MDS + Cc;
          double v = (double)accC[s][r] + (double)sM[base + (j - i) * (MDS - 1)];
          const int d0 = base - i * MDS - j;
          double tr = 0.0;
#pragma unroll
          for (int k = 0; k < ND; k++) tr += (double)sM[d0 + k * (MDS + 1)];
          accC[s][r] = (T)(i == j ? v + tr : v);
        }
      }
    }
    __syncthreads();  // M no longer needed: its space takes the node blocks
    // epilogue 3: the upper node blocks (a <= b; rows and columns are already in rank order) at their
    // stored offsets (ke_block_offset); the lower half of a diagonal block comes from its mirror entry
    T* sK = sL;
    if (active) {
#pragma unroll
      for (int s = 0; s < TPW; s++) {
        if (wv + WPE * s >= NTILE) continue;
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int R = 16 * tIv[s] + MT::row(l, r), Cc = 16 * tJv[s] + (l & 15);
          if (R >= NC || Cc >= NC || R > Cc) continue;
          const int i = R % ND, j = Cc % ND, a = R / ND, b = Cc / ND;
          const T v = accC[s][r];
          T* blk = sK + ke_block_offset(NN, ND, a, b);
          blk[i * ND + j] = v;
          if (a == b) blk[j * ND + i] = v;
        }
      }
      // padding values: the one between the halves (Hex20) and the tail
      constexpr int SR = ke_split_row(NN, ND);
      constexpr int PADS = SR < NN ? ke_block_offset(NN, ND, SR, SR) - 1 : -1;
      for (int i = EU + (PADS >= 0 ? 1 : 0) + te; i < EW; i += TPE) sK[i] = (T)0;
      if (PADS >= 0 && te == 0) sK[PADS] = (T)0;
    }
    __syncthreads();
    // epilogue 4: contiguous write of the element matrix in 16-byte pieces.  The trip count is a
    // compile-time constant so the store count is known where the next unit waits for its prefetch.
    if (active) {
      typedef T V __attribute__((ext_vector_type(16 / sizeof(T))));
      static_assert(EW * sizeof(T) % 16 == 0, "element matrix is a whole number of 16-byte pieces");
      constexpr int NV = EW * (int)sizeof(T) / 16, IT = (NV + TPE - 1) / TPE;
      V* dst = reinterpret_cast<V*>(ke + (u * EPW + el) * (int64_t)EW);
      const V* src = reinterpret_cast<const V*>(sK);
#pragma unroll
      for (int it = 0; it < IT; it++) {
        const int i = min(te + it * TPE, NV - 1);  // surplus lanes repeat the last piece (same value)
        dst[i] = src[i];
      }
    }
  }
}

constexpr size_t MFMA_LDS_TARGET = 56 * 1024;  // dynamic LDS per workgroup that the chunk size aims for
constexpr size_t MFMA_LDS_MAX = 64 * 1024;

template <int ND, int NN, int PREC>
int mfma_chunk(int nip, int axisym) {
  constexpr int KS = MfmaT<PREC>::KS;
  int ipc = (nip + KS - 1) / KS * KS;
  while (ipc > KS && mfma_layout<ND, NN, PREC>(nip, ipc, axisym).total * sizeof(double) > MFMA_LDS_TARGET) ipc -= KS;
  return ipc;
}

template <int ND, int NN, int PREC>
bool launch_mfma_kind(pm_fem* f, Group& g, const MatM& mp) {
  using S = MShape<ND, NN>;
  using T = typename MfmaT<PREC>::T;
  const int ax = f->axisym ? 1 : 0;
  // one lane per IP, and one prefetched state value per thread
  if (g.nip > 64 || (int64_t)S::EPW * S::SV * g.nip > (int64_t)S::NTHR) return false;
  const int ipc = mfma_chunk<ND, NN, PREC>(g.nip, ax);
  const size_t shm = mfma_layout<ND, NN, PREC>(g.nip, ipc, ax).total * sizeof(double);
  if (shm > MFMA_LDS_MAX) return false;
  const int64_t nunits = (g.nel + S::EPW - 1) / S::EPW;
  // resident workgroups per CU (registers and LDS), so that the persistent grid is one wave of workgroups
  int occ = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, reinterpret_cast<const void*>(&k_tangent_mfma<ND, NN, PREC>), S::NTHR, shm) != hipSuccess || occ < 1)
    occ = (int)std::max<int64_t>(1, f->lds_per_cu / (int64_t)(shm + 1024));
  const int64_t per_cu = std::max<int64_t>(1, std::min<int64_t>(occ, 32 / (S::WPE * S::EPW)));
  const int64_t nwg = std::min<int64_t>(nunits, per_cu * f->n_cu);
  const int64_t upw = (nunits + nwg - 1) / nwg;
  const int64_t grid = (nunits + upw - 1) / upw;
  if ((int64_t)S::EPW * g.nip * S::GW > 2 * (int64_t)S::NTHR) return false;  // the geometry prefetch roles
  if (!ke_range_ok(g.ke_base, g.nel, g.kew, packed_ke_values(NN, ND), f->n_ke_values)) {
    f->launch_refused = true;  // reported by check_device_error; never launched out of range
    return true;
  }
  T* ke;  // ke_base counts values
  if constexpr (PREC == 0) ke = f->d_ke + g.ke_base;
  else ke = f->d_ke32 + g.ke_base;
  hipLaunchKernelGGL((k_tangent_mfma<ND, NN, PREC>), dim3((unsigned)grid), dim3(S::NTHR), shm, f->stream, g.nel, g.nip, ipc, upw,
                     g.d_rank, g.d_geo, g.d_tab, g.d_state, ke, mp, ax, f->thickness, f->d_err,
                     (uint64_t)(&g - f->groups.data()) << kErrKeyShift);
  return true;
}

template <int ND, int NN>
bool launch_mfma_prec(pm_fem* f, Group& g, const MatM& mp, int prec) {
  if (prec == 0) return launch_mfma_kind<ND, NN, 0>(f, g, mp);
  if (prec == 1) return launch_mfma_kind<ND, NN, 1>(f, g, mp);
  return launch_mfma_kind<ND, NN, 2>(f, g, mp);
}

}  // namespace

// Launches the matrix-core tangent of one group; false when the kind or its IP count has no variant.
bool launch_tangent_mfma(pm_fem* f, Group& g, int prec) {
  if (g.nel == 0) return true;
  MatM mp;
  mp.model = g.param.model;
  mp.plane_stress = g.plane_stress ? 1 : 0;
  mp.G3 = 3.0 * g.G;
  mp.G2 = 2.0 * g.G;
  mp.G6G = 6.0 * g.G * g.G;
  mp.inv_d = 1.0 / (3.0 * g.G + g.param.hh);
  mp.K = g.K;
  mp.beta_e = g.K - mp.G2 / 3.0;
  const double E = g.param.young, nu = g.param.poisson;
  mp.a_ps = E / (1.0 + nu);
  mp.b_ps = E * nu / (1.0 - nu * nu);
  switch (g.kind) {
    case TRI3: return launch_mfma_prec<2, 3>(f, g, mp, prec);
    case TRI6: return launch_mfma_prec<2, 6>(f, g, mp, prec);
    case TRI10: return launch_mfma_prec<2, 10>(f, g, mp, prec);
    case TRI15: return launch_mfma_prec<2, 15>(f, g, mp, prec);
    case QUA4: return launch_mfma_prec<2, 4>(f, g, mp, prec);
    case QUA8: return launch_mfma_prec<2, 8>(f, g, mp, prec);
    case QUA9: return launch_mfma_prec<2, 9>(f, g, mp, prec);
    case TET4: return launch_mfma_prec<3, 4>(f, g, mp, prec);
    case TET10: return launch_mfma_prec<3, 10>(f, g, mp, prec);
    case HEX8: return launch_mfma_prec<3, 8>(f, g, mp, prec);
    case HEX20: return launch_mfma_prec<3, 20>(f, g, mp, prec);
    default: return false;
  }
}

}  // namespace pm
