// Device buffers of a pm_fem handle: allocation, upload of the element batches, patterns and
// gather maps, and release.
#include <algorithm>
#include <cstring>

#include "pm_internal.hpp"

namespace pm {

template <class T>
static int dalloc(pm_fem* f, T** p, int64_t n) {
  *p = nullptr;
  if (n <= 0) n = 1;
  hipError_t e = hipMalloc((void**)p, sizeof(T) * (size_t)n);
  if (e != hipSuccess) {
    f->err = std::string("device allocation failed (") + hipGetErrorString(e) + ")";
    return 1;
  }
  f->device_bytes += (int64_t)sizeof(T) * n;
  return 0;
}

template <class T>
static int upload(pm_fem* f, T* d, const T* h, int64_t n) {
  if (n <= 0) return 0;
  PM_HIP(hipMemcpyAsync(d, h, sizeof(T) * (size_t)n, hipMemcpyHostToDevice, f->stream));
  return 0;
}

int device_alloc(pm_fem* f) {
  PM_HIP(hipSetDevice(f->device));
  PM_HIP(hipStreamCreateWithFlags(&f->stream, hipStreamNonBlocking));
  {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, f->device) == hipSuccess && v > 0) f->n_cu = v;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, f->device) == hipSuccess && v > 0) f->lds_per_cu = v;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, f->device) == hipSuccess && v > 0) f->lds_per_block = v;
  }
  const int nd = f->ndim, ns = f->nsig;
  const int64_t b2 = (int64_t)nd * nd;
  if (dalloc(f, &f->d_X, f->npoint * nd) || upload(f, f->d_X, f->X.data(), f->npoint * nd)) return 1;
  // element batches
  for (auto& g : f->groups) {
    std::vector<int32_t> conn((size_t)g.nel * g.nnode);
    for (int64_t el = 0; el < g.nel; el++) {
      int64_t e = g.cells[el];
      std::copy(f->cpts.begin() + f->cptr[e], f->cpts.begin() + f->cptr[e + 1], conn.begin() + el * g.nnode);
    }
    Rule R;
    gauss_rule(g.kind, g.nip, R);
    std::vector<double> tab((size_t)g.nip * (1 + g.nnode + g.nnode * nd));
    double* w = tab.data();
    double* N = w + g.nip;
    double* dN = N + (size_t)g.nip * g.nnode;
    for (int q = 0; q < g.nip; q++) {
      w[q] = R.w[q];
      shape(g.kind, &R.x[3 * q], N + (size_t)q * g.nnode, dN + (size_t)q * g.nnode * nd);
    }
    const int64_t nipg = g.nel * g.nip;
    std::vector<uint8_t> rank((size_t)g.nel * g.nnode);
    for (int64_t el = 0; el < g.nel; el++) {
      int rk[32];
      cell_ranks(f, g.cells[el], rk);
      for (int m = 0; m < g.nnode; m++) rank[(size_t)el * g.nnode + m] = (uint8_t)rk[m];
    }
    if (dalloc(f, &g.d_conn, (int64_t)conn.size()) || upload(f, g.d_conn, conn.data(), (int64_t)conn.size())) return 1;
    if (dalloc(f, &g.d_rank, (int64_t)rank.size()) || upload(f, g.d_rank, rank.data(), (int64_t)rank.size())) return 1;
    if (dalloc(f, &g.d_tab, (int64_t)tab.size()) || upload(f, g.d_tab, tab.data(), (int64_t)tab.size())) return 1;
    if (dalloc(f, &g.d_state, nipg * (ns + 4)) || dalloc(f, &g.d_backup, nipg * (ns + 4))) return 1;
    if (dalloc(f, &g.d_geo, nipg * (nd * nd + 2))) return 1;
    launch_geometry(f, g);  // the mesh is fixed: J^-1 and det J of every IP are formed once
    PM_HIP(hipStreamSynchronize(f->stream));  // host vectors go out of scope
  }
  // patterns and maps
  // K is held as its upper node blocks; each row's lower blocks are listed with their stored mirrors
  const int64_t nnzl = (int64_t)f->h_lcol.size();
  if (dalloc(f, &f->d_urp, f->npoint + 1) || upload(f, f->d_urp, f->h_urp.data(), f->npoint + 1)) return 1;
  if (dalloc(f, &f->d_ucol, f->nnzu) || upload(f, f->d_ucol, f->h_ucol.data(), f->nnzu)) return 1;
  if (dalloc(f, &f->d_lrp, f->npoint + 1) || upload(f, f->d_lrp, f->h_lrp.data(), f->npoint + 1)) return 1;
  if (dalloc(f, &f->d_lcol, nnzl) || upload(f, f->d_lcol, f->h_lcol.data(), nnzl)) return 1;
  if (dalloc(f, &f->d_lidx, nnzl) || upload(f, f->d_lidx, f->h_lidx.data(), nnzl)) return 1;
  if (dalloc(f, &f->d_kmap_ptr, f->nnzu + 1) || upload(f, f->d_kmap_ptr, f->h_kmap_ptr.data(), f->nnzu + 1)) return 1;
  if (dalloc(f, &f->d_korder, (int64_t)f->h_korder.size()) || upload(f, f->d_korder, f->h_korder.data(), (int64_t)f->h_korder.size())) return 1;
  f->nkwave = (int64_t)f->h_kwave.size() - 1;
  if (dalloc(f, &f->d_kwave, (int64_t)f->h_kwave.size()) || upload(f, f->d_kwave, f->h_kwave.data(), (int64_t)f->h_kwave.size())) return 1;
  if (dalloc(f, &f->d_kmap, (int64_t)f->h_kmap.size()) || upload(f, f->d_kmap, f->h_kmap.data(), (int64_t)f->h_kmap.size())) return 1;
  if (dalloc(f, &f->d_fmap_ptr, f->npoint + 1) || upload(f, f->d_fmap_ptr, f->h_fmap_ptr.data(), f->npoint + 1)) return 1;
  if (dalloc(f, &f->d_fmap, (int64_t)f->h_fmap.size()) || upload(f, f->d_fmap, f->h_fmap.data(), (int64_t)f->h_fmap.size())) return 1;
  // K values: FP64 for FP64 tangents; FP32 only for reduced-precision tangents (formed, stored and
  // multiplied in FP32; readers that need FP64 widen them)
  if (f->cfg.tangent_precision == 0) {
    if (dalloc(f, &f->d_kval, f->nnzu * b2)) return 1;
    if (dalloc(f, &f->d_ke, f->n_ke_values)) return 1;
  } else {
    if (dalloc(f, &f->d_ke32, f->n_ke_values) || dalloc(f, &f->d_kval32, f->nnzu * b2)) return 1;
  }
  if (dalloc(f, &f->d_fe, f->n_fe)) return 1;
  // vectors
  const int64_t n = f->ndof, npr = std::max<int64_t>(f->nprescribed, 1);
  if (dalloc(f, &f->d_u, n) || dalloc(f, &f->d_ddu, n) || dalloc(f, &f->d_fint, n) || dalloc(f, &f->d_fext, n) ||
      dalloc(f, &f->d_rr, n) || dalloc(f, &f->d_mdu, n) || dalloc(f, &f->d_mdu0, n + npr) || dalloc(f, &f->d_lam, npr) ||
      dalloc(f, &f->d_y, npr) || dalloc(f, &f->d_cval, npr) || dalloc(f, &f->d_peq, npr) || dalloc(f, &f->d_free, n) ||
      dalloc(f, &f->d_pcg, 8 * n + 2 * npr) || dalloc(f, &f->d_binv, f->npoint * b2))
    return 1;
  const int64_t gn = (f->npoint + 255) / 256;
  if (dalloc(f, &f->d_red, std::max<int64_t>(4096 + 2 * std::min<int64_t>(gn, 2048), 1024) + 64) || dalloc(f, &f->d_scal, 32) || dalloc(f, &f->d_err, 4) ||
      dalloc(f, &f->d_work, (int64_t)kWorkCounters * kWorkStride)) return 1;
  std::vector<double> freev(n, 1.0);
  for (auto eq : f->prescribed_eq) freev[eq] = 0.0;
  if (upload(f, f->d_free, freev.data(), n)) return 1;
  if (upload(f, f->d_fext, f->f_ext.data(), n)) return 1;
  if (f->nprescribed && upload(f, f->d_peq, f->prescribed_eq.data(), f->nprescribed)) return 1;
  PM_HIP(hipMemsetAsync(f->d_u, 0, sizeof(double) * n, f->stream));
  PM_HIP(hipMemsetAsync(f->d_ddu, 0, sizeof(double) * n, f->stream));
  PM_HIP(hipMemsetAsync(f->d_lam, 0, sizeof(double) * npr, f->stream));
  PM_HIP(hipMemsetAsync(f->d_scal, 0, sizeof(double) * 32, f->stream));
  PM_HIP(hipStreamSynchronize(f->stream));
  reset_device_error(f);
  for (auto& g : f->groups) { launch_init_state(f, g); launch_copy(f, g.d_backup, g.d_state, g.nel * g.nip * (ns + 4)); }
  PM_HIP(hipStreamSynchronize(f->stream));
  // host copies of the maps are no longer needed
  std::vector<int32_t>().swap(f->h_kmap);
  std::vector<int32_t>().swap(f->h_kmap_ptr);
  std::vector<int32_t>().swap(f->h_korder);
  std::vector<int32_t>().swap(f->h_kwave);
  std::vector<int32_t>().swap(f->h_fmap);
  std::vector<int32_t>().swap(f->h_fmap_ptr);
  return 0;
}

// The multi-workgroup PCG streams a full-pattern copy of K, widened from the stored upper blocks at each
// solve: reading each lower block through its mirror costs more over a long solve than one widening.
// Allocated at the first such solve (systems solved by the single-workgroup PCG never need it).
int full_pattern_alloc(pm_fem* f) {
  if (f->d_kfull || f->d_kfull32) return 0;  // allocated last: set only once every piece is in place
  const int64_t b2 = (int64_t)f->bs * f->bs;
  const bool ok = !dalloc(f, &f->d_frp, f->npoint + 1) && !upload(f, f->d_frp, f->h_browptr.data(), f->npoint + 1) &&
                  !dalloc(f, &f->d_fcol, f->nnzb) && !upload(f, f->d_fcol, f->h_bcols.data(), f->nnzb) &&
                  !dalloc(f, &f->d_f2u, f->nnzb) && !upload(f, f->d_f2u, f->h_full2up.data(), f->nnzb) &&
                  !(f->cfg.tangent_precision == 0 ? dalloc(f, &f->d_kfull, f->nnzb * b2) : dalloc(f, &f->d_kfull32, f->nnzb * b2));
  if (ok) return 0;
  // a failed allocation leaves nothing behind (and is not counted), so a later solve tries again from scratch
  (void)hipStreamSynchronize(f->stream);  // uploads into these buffers end before they are freed
  const int64_t vsz = f->cfg.tangent_precision == 0 ? (int64_t)sizeof(double) : (int64_t)sizeof(float);
  const std::pair<void*, int64_t> made[] = {{f->d_frp, 4 * (f->npoint + 1)}, {f->d_fcol, 4 * f->nnzb}, {f->d_f2u, 4 * f->nnzb},
                                            {f->d_kfull, vsz * f->nnzb * b2}, {f->d_kfull32, vsz * f->nnzb * b2}};
  for (const auto& m : made)
    if (m.first) { (void)hipFree(m.first); f->device_bytes -= std::max<int64_t>(m.second, (int64_t)sizeof(int32_t)); }
  f->d_frp = f->d_fcol = f->d_f2u = nullptr;
  f->d_kfull = nullptr;
  f->d_kfull32 = nullptr;
  return 1;
}

int arc_alloc(pm_fem* f) {
  if (f->d_arc) return 0;
  return dalloc(f, &f->d_arc, 4 * f->ndof + 4 * std::max<int64_t>(f->nprescribed, 1));
}

void device_free(pm_fem* f) {
  release_graphs(f);
  auto fr = [](void* p) { if (p) (void)hipFree(p); };
  for (auto& g : f->groups) { fr(g.d_conn); fr(g.d_rank); g.d_rank = nullptr; fr(g.d_tab); fr(g.d_state); fr(g.d_backup); fr(g.d_geo); g.d_conn = nullptr; g.d_tab = g.d_state = g.d_backup = g.d_geo = nullptr; }
  fr(f->d_X); fr(f->d_urp); fr(f->d_ucol); fr(f->d_lrp); fr(f->d_lcol); fr(f->d_lidx); fr(f->d_kval); fr(f->d_kval32); fr(f->d_ke); fr(f->d_ke32);
  fr(f->d_kmap_ptr); fr(f->d_kmap); fr(f->d_korder); fr(f->d_kwave); fr(f->d_fe); fr(f->d_fmap_ptr); fr(f->d_fmap);
  fr(f->d_u); fr(f->d_ddu); fr(f->d_fint); fr(f->d_fext); fr(f->d_rr); fr(f->d_mdu); fr(f->d_mdu0); fr(f->d_lam);
  fr(f->d_y); fr(f->d_cval); fr(f->d_peq); fr(f->d_free); fr(f->d_pcg); fr(f->d_binv); fr(f->d_red); fr(f->d_scal); fr(f->d_err); fr(f->d_work);
  fr(f->d_arc); f->d_arc = nullptr;
  fr(f->d_frp); fr(f->d_fcol); fr(f->d_f2u); fr(f->d_kfull); fr(f->d_kfull32);
  f->d_frp = f->d_fcol = f->d_f2u = nullptr; f->d_kfull = nullptr; f->d_kfull32 = nullptr;
  if (f->stream) (void)hipStreamDestroy(f->stream);
  f->stream = nullptr;
  f->d_X = nullptr; f->d_urp = f->d_ucol = f->d_lrp = f->d_lcol = f->d_lidx = nullptr; f->d_kval = nullptr; f->d_kval32 = nullptr; f->d_ke = nullptr; f->d_ke32 = nullptr;
  f->d_kmap_ptr = f->d_kmap = nullptr; f->d_korder = nullptr; f->d_kwave = nullptr; f->d_fe = nullptr; f->d_fmap_ptr = f->d_fmap = nullptr;
  f->d_u = f->d_ddu = f->d_fint = f->d_fext = f->d_rr = f->d_mdu = f->d_mdu0 = f->d_lam = nullptr;
  f->d_y = f->d_cval = nullptr; f->d_peq = nullptr; f->d_free = f->d_pcg = f->d_binv = f->d_red = f->d_scal = nullptr; f->d_err = nullptr; f->d_work = nullptr;
}

}  // namespace pm
