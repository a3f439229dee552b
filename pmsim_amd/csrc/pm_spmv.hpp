// Row of K x for the node-block tangent held as its upper triangle.
//
// K is symmetric and is stored as the node blocks (a, b) with b >= a only (urp / ucol / vals).  The
// blocks below the diagonal of row a are read as the transposes of their mirrors: lrp / lcol list, in
// ascending column, the columns b < a of row a, and lidx the index of the stored block (b, a).  A row
// takes its lower blocks first, then its upper ones, so the columns come in ascending order with the
// same values as the full matrix: the products and their summation order are those of a full-CSR row.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace pm {

template <typename VT>
struct KSym {
  const int32_t* urp;   // npoint + 1
  const int32_t* ucol;  // upper blocks
  const int32_t* lrp;   // npoint + 1
  const int32_t* lcol;  // lower blocks of each row (ascending column)
  const int32_t* lidx;  // stored upper block of each lower entry
  const VT* vals;       // upper blocks, BS x BS row-major each
};

// component i of (K x)[a]
template <int BS, typename VT>
__device__ __forceinline__ double ksym_row(const KSym<VT>& K, int64_t a, int i, const double* __restrict__ x) {
  double s = 0.0;
  for (int32_t k = K.lrp[a]; k < K.lrp[a + 1]; k++) {
    const int64_t b = K.lcol[k];
    const VT* v = K.vals + (int64_t)K.lidx[k] * BS * BS + i;  // K_ab[i][j] = K_ba[j][i]
#pragma unroll
    for (int j = 0; j < BS; j++) s += (double)v[j * BS] * x[b * BS + j];
  }
  for (int32_t k = K.urp[a]; k < K.urp[a + 1]; k++) {
    const int64_t b = K.ucol[k];
    const VT* v = K.vals + (int64_t)k * BS * BS + i * BS;
#pragma unroll
    for (int j = 0; j < BS; j++) s += (double)v[j] * x[b * BS + j];
  }
  return s;
}

// Row a of K x over the full-pattern copy of K (row pointers rp, columns cols, BS x BS row-major blocks),
// taken by a group of lpr lanes (lpr a power of two <= 64, the group aligned in the wave): lane l
// multiplies the row's blocks l, l + lpr, ... (consecutive lanes read consecutive blocks), then a
// butterfly over the group leaves the row's sums y in every lane of the group.  Every lane of the group
// must call it for the same row.  The same arithmetic whatever the caller, so the single-request and the
// batched PCG give the same bits.
template <int BS, typename VT>
__device__ __forceinline__ void full_row_lanes(const int32_t* __restrict__ rp, const int32_t* __restrict__ cols,
                                               const VT* __restrict__ vals, int64_t a, int lane, int lpr,
                                               const double* __restrict__ x, double* y) {
#pragma unroll
  for (int i = 0; i < BS; i++) y[i] = 0.0;
  for (int32_t k = rp[a] + lane; k < rp[a + 1]; k += lpr) {
    const int64_t b = cols[k];
    const VT* v = vals + (int64_t)k * BS * BS;
    double xb[BS];
#pragma unroll
    for (int j = 0; j < BS; j++) xb[j] = x[b * BS + j];
#pragma unroll
    for (int i = 0; i < BS; i++)
#pragma unroll
      for (int j = 0; j < BS; j++) y[i] += (double)v[i * BS + j] * xb[j];
  }
  // lane ^ off stays inside the aligned group for off < lpr
  for (int off = lpr >> 1; off > 0; off >>= 1)
#pragma unroll
    for (int i = 0; i < BS; i++) y[i] += __shfl_xor(y[i], off);
}

// full_row_lanes for two rows a0 and a1 at once (twice the loads in flight); each row's sums are formed
// exactly as full_row_lanes forms them
template <int BS, typename VT>
__device__ __forceinline__ void full_row_lanes2(const int32_t* __restrict__ rp, const int32_t* __restrict__ cols,
                                                const VT* __restrict__ vals, int64_t a0, int64_t a1, int lane, int lpr,
                                                const double* __restrict__ x, double* y0, double* y1) {
#pragma unroll
  for (int i = 0; i < BS; i++) { y0[i] = 0.0; y1[i] = 0.0; }
  int32_t k0 = rp[a0] + lane, k1 = rp[a1] + lane;
  const int32_t e0 = rp[a0 + 1], e1 = rp[a1 + 1];
  while (k0 < e0 || k1 < e1) {
    if (k0 < e0) {
      const int64_t b = cols[k0];
      const VT* v = vals + (int64_t)k0 * BS * BS;
      double xb[BS];
#pragma unroll
      for (int j = 0; j < BS; j++) xb[j] = x[b * BS + j];
#pragma unroll
      for (int i = 0; i < BS; i++)
#pragma unroll
        for (int j = 0; j < BS; j++) y0[i] += (double)v[i * BS + j] * xb[j];
      k0 += lpr;
    }
    if (k1 < e1) {
      const int64_t b = cols[k1];
      const VT* v = vals + (int64_t)k1 * BS * BS;
      double xb[BS];
#pragma unroll
      for (int j = 0; j < BS; j++) xb[j] = x[b * BS + j];
#pragma unroll
      for (int i = 0; i < BS; i++)
#pragma unroll
        for (int j = 0; j < BS; j++) y1[i] += (double)v[i * BS + j] * xb[j];
      k1 += lpr;
    }
  }
  for (int off = lpr >> 1; off > 0; off >>= 1)
#pragma unroll
    for (int i = 0; i < BS; i++) {
      y0[i] += __shfl_xor(y0[i], off);
      y1[i] += __shfl_xor(y1[i], off);
    }
}

// lanes per block row for full_row_lanes, from the mean row length: 32 for Hex20 (57 blocks per row; 64
// lanes, one block each, measured 8 % slower: a longer butterfly), 8 for 2D quadratic meshes (about 16:
// two blocks per lane measured faster than one on small and large systems alike)
inline int lanes_per_row(int64_t nnzb, int64_t np) {
  const int64_t mean_row = (nnzb + np - 1) / (np > 0 ? np : 1);
  return mean_row > 16 ? 32 : 8;
}

// workgroups of tpb threads for an SpMV of np block rows at lpr lanes per row: one row per lane group
// where that fits in 4096 workgroups (short latency chains on small systems), else a grid-stride loop
inline int spmv_workgroups(int64_t np, int lpr, int tpb) {
  const int64_t w = (np * lpr + tpb - 1) / tpb;
  return (int)(w < 1 ? 1 : w > 4096 ? 4096 : w);
}

}  // namespace pm
