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

}  // namespace pm
