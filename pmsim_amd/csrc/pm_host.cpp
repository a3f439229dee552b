// Host-side setup: element batches, node-block CSR pattern, deterministic scatter maps,
// external forces, and the host direct solver used as fallback and as the check of the iterative solves.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <numeric>
#include <string>

#include "pm_internal.hpp"

namespace pm {

// an f64 as Rust's {:?} prints it: the shortest form that reads back to the value, ".0" on integral values
static std::string rust_debug_f64(double v) {
  char b[48];
  for (int p = 1; p <= 17; p++) {
    std::snprintf(b, sizeof b, "%.*g", p, v);
    if (std::strtod(b, nullptr) == v) break;
  }
  std::string s(b);
  const size_t e = s.find('e');
  if (e != std::string::npos) {  // exponent without '+' or leading zeros
    std::string m = s.substr(0, e), x = s.substr(e + 1);
    const bool neg = !x.empty() && x[0] == '-';
    if (!x.empty() && (x[0] == '-' || x[0] == '+')) x = x.substr(1);
    while (x.size() > 1 && x[0] == '0') x = x.substr(1);
    return m + "e" + (neg ? "-" : "") + x;
  }
  if (s.find_first_of(".n") == std::string::npos) s += ".0";  // 'n': nan, inf
  return s;
}

// Element batches: one per (kind, attribute); cells keep ascending global order.
int build_groups(pm_fem* f) {
  f->nsig = f->ndim == 2 ? 4 : 6;
  f->ndof = (int64_t)f->ndim * f->npoint;
  std::map<std::pair<int, int>, int> key2g;
  f->cell_group.assign(f->ncell, -1);
  f->cell_local.assign(f->ncell, 0);
  for (int64_t e = 0; e < f->ncell; e++) {
    int k = f->kind[e];
    if (k < 0 || k >= NKIND) { f->err = "invalid cell kind"; return 1; }
    if (kind_ndim(k) != f->ndim) { f->err = "cell kind does not match ndim"; return 1; }
    if (f->cptr[e + 1] - f->cptr[e] != kind_nnode(k)) { f->err = "wrong number of points for cell kind"; return 1; }
    auto key = std::make_pair(k, (int)f->attr[e]);
    auto it = key2g.find(key);
    if (it == key2g.end()) {
      auto pit = f->params.find(f->attr[e]);
      if (pit == f->params.end()) { f->err = "cannot find attribute in the element map"; return 1; }
      Group g;
      g.kind = k; g.ndim = f->ndim; g.nnode = kind_nnode(k); g.attribute = f->attr[e]; g.param = pit->second;
      int ng = g.param.ngauss > 0 ? g.param.ngauss : kind_default_ngauss(k);
      Rule r;
      if (!gauss_rule(k, ng, r)) { f->err = "requested number of integration points is not available"; return 1; }
      g.nip = r.n();
      g.kew = packed_ke_values(g.nnode, g.ndim);
      const ParamSolid& p = g.param;
      // ModelStressStrain::new (model_stress_strain.rs:52-72): the attribute's settings are validated first
      // (settings.rs:55-66); von Mises under general plasticity is the Elastoplastic model
      const auto sit = f->model_settings.find(g.attribute);
      const ModelSettings ms = sit != f->model_settings.end() ? sit->second : ModelSettings();
      std::string bad;
      if (ms.nle_beta < 0.0) bad = "nle_beta = " + rust_debug_f64(ms.nle_beta) + " is incorrect; it must be ≥ 0.0";
      else if (ms.gp_interp_nn_max < 1) bad = "gp_interp_nn_max = " + std::to_string(ms.gp_interp_nn_max) + " is incorrect; it must be ≥ 1";
      if (!bad.empty()) {
        std::printf("ERROR: %s\n", bad.c_str());  // settings.rs prints it before the allocation fails
        std::fflush(stdout);
        f->err = "cannot allocate ModelStressStrain because settings.validate() failed";
        return 1;
      }
      g.general_plasticity = p.model == 1 && ms.general_plasticity;
      if (p.model == 1) {
        if (f->plane_stress) { f->err = "von Mises model does not work in plane-stress"; return 1; }
        if (p.z_ini <= 1e-6) { f->err = "von Mises initial size of the yield surface must > 1e-6"; return 1; }
      }
      g.K = p.young / (3.0 * (1.0 - 2.0 * p.poisson));
      g.G = p.young / (2.0 * (1.0 + p.poisson));
      g.plane_stress = f->plane_stress;
      it = key2g.emplace(key, (int)f->groups.size()).first;
      f->groups.push_back(g);
    }
    Group& g = f->groups[it->second];
    f->cell_group[e] = it->second;
    f->cell_local[e] = (int64_t)g.cells.size();
    g.cells.push_back(e);
  }
  // IP numbering in cell order
  f->cell_ip.assign(f->ncell + 1, 0);
  for (int64_t e = 0; e < f->ncell; e++) f->cell_ip[e + 1] = f->cell_ip[e] + f->groups[f->cell_group[e]].nip;
  f->nip = f->cell_ip[f->ncell];
  int64_t fe = 0, ke = 0;
  for (auto& g : f->groups) {
    g.nel = (int64_t)g.cells.size();
    g.fe_base = fe; fe += g.nel * g.nnode * g.ndim;
    g.ke_base = ke; ke += g.nel * g.kew;
  }
  f->n_fe = fe;
  f->n_ke_values = ke;
  // every point must belong to a cell (pure solid DOF map)
  std::vector<uint8_t> used(f->npoint, 0);
  for (auto p : f->cpts) {
    if (p < 0 || p >= f->npoint) { f->err = "cell point id out of range"; return 1; }
    used[p] = 1;
  }
  for (auto u : used) if (!u) { f->err = "all points must belong to at least one cell"; return 1; }
  // prescribed equations
  f->nprescribed = (int64_t)f->ess_point.size();
  f->prescribed_eq.resize(f->nprescribed);
  f->is_prescribed.assign(f->ndof, 0);
  for (int64_t i = 0; i < f->nprescribed; i++) {
    if (f->ess_point[i] < 0 || f->ess_point[i] >= f->npoint || f->ess_dof[i] < 0 || f->ess_dof[i] >= f->ndim) {
      f->err = "cannot find (PointId, Dof) in the DOF map"; return 1;
    }
    int64_t eq = (int64_t)f->ndim * f->ess_point[i] + f->ess_dof[i];
    f->prescribed_eq[i] = eq;
    f->is_prescribed[eq] = 1;
  }
  return 0;
}

// rank of each local node of cell e among the cell's global point ids (ascending)
void cell_ranks(const pm_fem* f, int64_t e, int* rk) {
  const int32_t* pts = &f->cpts[f->cptr[e]];
  const int n = (int)(f->cptr[e + 1] - f->cptr[e]);
  for (int a = 0; a < n; a++) {
    int r = 0;
    for (int b = 0; b < n; b++) r += (pts[b] < pts[a]) ? 1 : 0;
    rk[a] = r;
  }
}

// Node-block CSR pattern of K and the two deterministic gather maps:
//  * K: for each block (a,b), the element blocks that contribute, in ascending global cell order
//    (pmsim's COO insertion order, src/base/assembly.rs:124-135; duplicates summed in that order);
//  * R: for each node, the element force vectors that contribute, in ascending global cell order.
int build_patterns(pm_fem* f) {
  const int64_t np = f->npoint;
  f->bs = f->ndim;
  // node -> cells (ascending)
  std::vector<int64_t> nptr(np + 1, 0);
  for (auto p : f->cpts) nptr[p + 1]++;
  for (int64_t i = 0; i < np; i++) nptr[i + 1] += nptr[i];
  std::vector<int64_t> ncells(nptr[np]);
  std::vector<int32_t> nloc(nptr[np]);
  {
    std::vector<int64_t> fill(nptr.begin(), nptr.end() - 1);
    for (int64_t e = 0; e < f->ncell; e++)
      for (int64_t k = f->cptr[e]; k < f->cptr[e + 1]; k++) {
        int32_t p = f->cpts[k];
        ncells[fill[p]] = e;
        nloc[fill[p]] = (int32_t)(k - f->cptr[e]);
        fill[p]++;
      }
  }
  // block rows
  f->h_browptr.assign(np + 1, 0);
  std::vector<std::vector<int32_t>> rows(np);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t a = 0; a < np; a++) {
    std::vector<int32_t>& v = rows[a];
    for (int64_t k = nptr[a]; k < nptr[a + 1]; k++) {
      int64_t e = ncells[k];
      for (int64_t q = f->cptr[e]; q < f->cptr[e + 1]; q++) v.push_back(f->cpts[q]);
    }
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
  }
  int64_t nnzb = 0;
  for (int64_t a = 0; a < np; a++) { nnzb += (int64_t)rows[a].size(); }
  if (nnzb >= (int64_t)INT32_MAX) { f->err = "too many nonzero blocks for int32 indices"; return 1; }
  f->nnzb = nnzb;
  f->h_bcols.resize(nnzb);
  for (int64_t a = 0; a < np; a++) {
    f->h_browptr[a + 1] = f->h_browptr[a] + (int32_t)rows[a].size();
    std::copy(rows[a].begin(), rows[a].end(), f->h_bcols.begin() + f->h_browptr[a]);
    std::vector<int32_t>().swap(rows[a]);
  }
  // upper pattern (b >= a), full -> upper map and the lower blocks of each row (the pattern is symmetric)
  f->h_urp.assign(np + 1, 0);
  for (int64_t a = 0; a < np; a++) {
    int32_t nu = 0;
    for (int32_t k = f->h_browptr[a]; k < f->h_browptr[a + 1]; k++) nu += f->h_bcols[k] >= a ? 1 : 0;
    f->h_urp[a + 1] = f->h_urp[a] + nu;
  }
  const int64_t nnzu = f->h_urp[np];
  if (2 * nnzu != nnzb + np) { f->err = "internal error: the block pattern is not symmetric"; return 1; }
  f->nnzu = nnzu;
  f->h_ucol.resize(nnzu);
  for (int64_t a = 0; a < np; a++) {
    int32_t u = f->h_urp[a];
    for (int32_t k = f->h_browptr[a]; k < f->h_browptr[a + 1]; k++)
      if (f->h_bcols[k] >= a) f->h_ucol[u++] = f->h_bcols[k];
  }
  for (int64_t a = 0; a < np; a++)  // block Jacobi reads the diagonal as the first stored block of a row
    if (f->h_urp[a] == f->h_urp[a + 1] || f->h_ucol[f->h_urp[a]] != a) { f->err = "internal error: a row without its diagonal block"; return 1; }
  auto upper_index = [&](int32_t a, int32_t b) -> int64_t {  // stored block (a, b), a <= b
    const int32_t* lo = f->h_ucol.data() + f->h_urp[a];
    const int32_t* hi = f->h_ucol.data() + f->h_urp[a + 1];
    const int32_t* it = std::lower_bound(lo, hi, b);
    return (it != hi && *it == b) ? (int64_t)(it - f->h_ucol.data()) : -1;
  };
  f->h_full2up.resize(nnzb);
  f->h_lrp.assign(np + 1, 0);
  f->h_lcol.clear();
  f->h_lidx.clear();
  f->h_lcol.reserve(nnzb - nnzu);
  f->h_lidx.reserve(nnzb - nnzu);
  for (int64_t a = 0; a < np; a++) {
    for (int32_t k = f->h_browptr[a]; k < f->h_browptr[a + 1]; k++) {
      const int32_t b = f->h_bcols[k];
      const int64_t u = b >= a ? upper_index((int32_t)a, b) : upper_index(b, (int32_t)a);
      if (u < 0) { f->err = "internal error: the block pattern is not symmetric"; return 1; }
      f->h_full2up[k] = b >= a ? (int32_t)u : (int32_t)((uint32_t)u | 0x80000000u);
      if (b < a) { f->h_lcol.push_back(b); f->h_lidx.push_back((int32_t)u); }
    }
    f->h_lrp[a + 1] = (int32_t)f->h_lcol.size();
  }
  // K gather map over the stored (upper) blocks: element node pairs (la, lb) with point(la) <= point(lb).
  // Element matrices hold their nodes in rank order (ascending point id), so such a pair is the stored
  // element block (rank[la], rank[lb]) with rank[la] <= rank[lb]: every read is direct.
  std::vector<int32_t> kptr(nnzu + 1, 0);
  int64_t total = 0;
  for (int64_t e = 0; e < f->ncell; e++) { int64_t n = f->cptr[e + 1] - f->cptr[e]; total += n * (n + 1) / 2; }
  if (f->n_ke_values >= (int64_t)INT32_MAX || total >= (int64_t)INT32_MAX) { f->err = "mesh too large for int32 scatter maps"; return 1; }
  std::vector<int32_t> pairblk(total);
  std::vector<int64_t> pbase(f->ncell + 1, 0);
  for (int64_t e = 0; e < f->ncell; e++) { int64_t n = f->cptr[e + 1] - f->cptr[e]; pbase[e + 1] = pbase[e] + n * (n + 1) / 2; }
  bool bad_pair = false;
#pragma omp parallel for schedule(static) reduction(|| : bad_pair)
  for (int64_t e = 0; e < f->ncell; e++) {
    int64_t n = f->cptr[e + 1] - f->cptr[e], w = pbase[e];
    const int32_t* pts = &f->cpts[f->cptr[e]];
    for (int64_t la = 0; la < n; la++)
      for (int64_t lb = 0; lb < n; lb++) {
        if (pts[la] > pts[lb]) continue;
        const int64_t u = upper_index(pts[la], pts[lb]);
        if (u < 0) bad_pair = true;
        pairblk[w++] = (int32_t)u;
      }
  }
  if (bad_pair) { f->err = "internal error: element pair outside the block pattern"; return 1; }
  for (int64_t i = 0; i < total; i++) kptr[pairblk[i] + 1]++;
  for (int64_t i = 0; i < nnzu; i++) kptr[i + 1] += kptr[i];
  std::vector<int32_t> kmap(total);
  {
    std::vector<int32_t> fill(kptr.begin(), kptr.end() - 1);
    for (int64_t e = 0; e < f->ncell; e++) {  // ascending cells => ascending contributions
      const Group& g = f->groups[f->cell_group[e]];
      int64_t n = g.nnode, el = f->cell_local[e], w = pbase[e];
      const int32_t* pts = &f->cpts[f->cptr[e]];
      int rk[32];
      cell_ranks(f, e, rk);
      for (int64_t la = 0; la < n; la++)
        for (int64_t lb = 0; lb < n; lb++) {
          if (pts[la] > pts[lb]) continue;
          const int ra = rk[la], rb = rk[lb];
          if (ra > rb) { f->err = "internal error: element ranks out of order"; return 1; }
          const int64_t off = g.ke_base + el * g.kew + (int64_t)ke_block_offset((int)n, g.ndim, ra, rb);
          kmap[fill[pairblk[w++]]++] = (int32_t)off;
        }
    }
  }
  // every map entry must address a whole block inside the element-matrix buffer
  for (int32_t m : kmap) {
    const int64_t off = (int64_t)((uint32_t)m & 0x7fffffffu);
    if (off + (int64_t)f->ndim * f->ndim > f->n_ke_values) { f->err = "internal error: K gather map out of range"; return 1; }
  }
  // Processing order of the gather's chunks (GATHER_CHUNK consecutive CSR blocks, one workgroup each): by the
  // smallest element-matrix offset a chunk reads, so chunks of different node layers that read the same
  // elements run close together and find those matrices in L2.  The sums do not depend on the order.
  {
    const int64_t nchunk = (nnzu + GATHER_CHUNK - 1) / GATHER_CHUNK;
    std::vector<int32_t> key(nchunk, INT32_MAX);
    for (int64_t c = 0; c < nchunk; c++)
      for (int64_t b = c * GATHER_CHUNK; b < std::min<int64_t>(nnzu, (c + 1) * GATHER_CHUNK); b++)
        for (int32_t k = kptr[b]; k < kptr[b + 1]; k++) key[c] = std::min(key[c], (int32_t)((uint32_t)kmap[k] & 0x7fffffffu));
    f->h_korder.resize(nchunk);
    std::iota(f->h_korder.begin(), f->h_korder.end(), 0);
    std::stable_sort(f->h_korder.begin(), f->h_korder.end(), [&](int32_t a, int32_t b) { return key[a] < key[b]; });
  }
  // Wavefront ranges of the balanced K gather: whole CSR blocks, in order, packed greedily so that a range
  // holds at most 64 contributions and 64 blocks (one lane per contribution, one owner lane per block); a
  // block with more than 64 contributions has a range of its own, taken in rounds of 64.
  {
    std::vector<int32_t> kw(1, 0);
    int64_t b = 0;
    while (b < nnzu) {
      const int64_t start = b++;
      if (kptr[start + 1] - kptr[start] <= 64)
        while (b < nnzu && b - start < 64 && kptr[b + 1] - kptr[start] <= 64) b++;
      kw.push_back((int32_t)b);
    }
    f->h_kwave.swap(kw);
  }
  // upload happens in device_alloc; keep host copies in temporaries stored on the handle
  f->h_kmap_ptr.swap(kptr);
  f->h_kmap.swap(kmap);
  // R gather map
  std::vector<int32_t> fptr(np + 1, 0), fmap(nptr[np]);
  for (int64_t a = 0; a < np; a++) {
    fptr[a + 1] = (int32_t)nptr[a + 1];
    for (int64_t k = nptr[a]; k < nptr[a + 1]; k++) {
      int64_t e = ncells[k];
      const Group& g = f->groups[f->cell_group[e]];
      int64_t off = g.fe_base / g.ndim + f->cell_local[e] * g.nnode + nloc[k];  // in ndim-vectors
      fmap[k] = (int32_t)off;
    }
  }
  f->h_fmap_ptr.swap(fptr);
  f->h_fmap.swap(fmap);
  return 0;
}

// External forces at time station k: concentrated loads on every equation (bc_concentrated.rs:46-54)
// and normal distributed loads (bc_distributed.rs:166-174: f = int N qn n dGamma with n the outward
// normal) on the equations that are not ignored, i.e. not prescribed under the reduced system method
// (solver_data.rs:97-107), each with its value at that station (pmsim evaluates value = f(t) at the
// step's time, bc_distributed.rs:161-164).  The prescribed rows of the reduced system are overwritten
// anyway, but F.F of the arc-length constraint sees them.
void assemble_f_ext(pm_fem* f, int64_t k) {
  const bool rsm = !f->cfg.lagrange_mult_method && f->is_prescribed.size() == (size_t)f->ndof;
  f->f_ext.assign(f->ndof, 0.0);
  const int64_t kp = std::min<int64_t>(std::max<int64_t>(k, 0), f->pl_nstation - 1);
  for (size_t i = 0; i < f->pl_point.size(); i++)
    f->f_ext[(int64_t)f->ndim * f->pl_point[i] + f->pl_dof[i]] += f->pl_table[i * f->pl_nstation + kp];
  const int nd = f->ndim;
  for (const Facet& F : f->facets) {
    int nn = (int)F.pts.size();
    const double value = F.values[std::min<int64_t>(std::max<int64_t>(k, 0), (int64_t)F.values.size() - 1)];
    Rule R;
    bool tri = (F.kind == TRI3 || F.kind == TRI6);
    if (F.kind >= LIN2) gauss_rule(F.kind, std::max(3, std::min(4, nn)), R);
    else if (tri) gauss_rule(F.kind, 6, R);
    else gauss_rule(QUA9, 9, R);
    // owner cell centroid and facet centroid (outward orientation)
    double cc[3] = {0, 0, 0}, fc[3] = {0, 0, 0};
    if (F.cell >= 0) {
      int64_t cn = f->cptr[F.cell + 1] - f->cptr[F.cell];
      for (int64_t k = 0; k < cn; k++) for (int i = 0; i < nd; i++) cc[i] += f->X[(int64_t)f->cpts[f->cptr[F.cell] + k] * nd + i] / cn;
    }
    for (int k = 0; k < nn; k++) for (int i = 0; i < nd; i++) fc[i] += f->X[(int64_t)F.pts[k] * nd + i] / nn;
    const int fd = kind_ndim(F.kind);
    for (int q = 0; q < R.n(); q++) {
      double N[20], dN[40];
      shape(F.kind, &R.x[3 * q], N, dN);
      // n: the facet's normal scaled by its measure (2D edge: (dy, -dx); 3D face: a x b); 3D edge: the tangent
      double n[3] = {0, 0, 0};
      if (nd == 2) {
        double dx = 0, dy = 0;
        for (int m = 0; m < nn; m++) { dx += f->X[(int64_t)F.pts[m] * 2] * dN[m]; dy += f->X[(int64_t)F.pts[m] * 2 + 1] * dN[m]; }
        n[0] = dy; n[1] = -dx;
      } else if (fd == 1) {
        for (int m = 0; m < nn; m++) for (int i = 0; i < 3; i++) n[i] += f->X[(int64_t)F.pts[m] * 3 + i] * dN[m];
      } else {
        double a[3] = {0, 0, 0}, b[3] = {0, 0, 0};
        for (int m = 0; m < nn; m++) for (int i = 0; i < 3; i++) {
          a[i] += f->X[(int64_t)F.pts[m] * 3 + i] * dN[m * fd];
          b[i] += f->X[(int64_t)F.pts[m] * 3 + i] * dN[m * fd + 1];
        }
        n[0] = a[1] * b[2] - a[2] * b[1]; n[1] = a[2] * b[0] - a[0] * b[2]; n[2] = a[0] * b[1] - a[1] * b[0];
      }
      double w = R.w[q] * (nd == 2 ? f->thickness : 1.0);
      if (nd == 2 && f->axisym) {
        double r = 0;
        for (int m = 0; m < nn; m++) r += N[m] * f->X[(int64_t)F.pts[m] * 2];
        w *= r;
      }
      // Qn: value times the outward normal (bc_distributed.rs:182-187); Qx/Qy/Qz: value in one direction per
      // unit facet measure |n| (bc_distributed.rs:195-231)
      double v[3] = {0, 0, 0};
      if (F.nbc == 0) {
        double dot = 0;
        for (int i = 0; i < nd; i++) dot += n[i] * (fc[i] - cc[i]);
        const double sgn = dot < 0 ? -1.0 : 1.0;
        for (int i = 0; i < nd; i++) v[i] = value * sgn * n[i];
      } else {
        v[F.nbc - 1] = value * std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
      }
      for (int m = 0; m < nn; m++)
        for (int i = 0; i < nd; i++) {
          const int64_t eq = (int64_t)F.pts[m] * nd + i;
          if (rsm && f->is_prescribed[eq]) continue;  // ignored equations (solver_data.rs:103)
          f->f_ext[eq] += N[m] * v[i] * w;
        }
    }
  }
  f->ftf = 0.0;
  for (double v : f->f_ext) f->ftf += v * v;
}

// ----------------------------------------------------------------------------------------------
// host direct solver: envelope (skyline) Cholesky of K_ff after a reverse Cuthill-McKee ordering
// ----------------------------------------------------------------------------------------------
// K values in the full block pattern (h_browptr / h_bcols), as FP64 numbers: the stored upper blocks are
// copied down (FP32 values widened) and each lower block is the transpose of its stored mirror.
int copy_kval_to_host(const pm_fem* f, std::vector<double>& kv) {
  const int nd = f->ndim, b2 = nd * nd;
  const size_t nu = (size_t)f->nnzu * b2;
  std::vector<double> up(nu);
  if (f->d_kval) {
    if (hipMemcpyAsync(up.data(), f->d_kval, nu * sizeof(double), hipMemcpyDeviceToHost, f->stream) != hipSuccess ||
        hipStreamSynchronize(f->stream) != hipSuccess)
      return 1;
  } else {
    std::vector<float> k32(nu);
    if (hipMemcpyAsync(k32.data(), f->d_kval32, nu * sizeof(float), hipMemcpyDeviceToHost, f->stream) != hipSuccess ||
        hipStreamSynchronize(f->stream) != hipSuccess)
      return 1;
    for (size_t i = 0; i < nu; i++) up[i] = (double)k32[i];
  }
  kv.resize((size_t)f->nnzb * b2);
  for (int64_t k = 0; k < f->nnzb; k++) {
    const uint32_t m = (uint32_t)f->h_full2up[k];
    const double* src = up.data() + (size_t)(m & 0x7fffffffu) * b2;
    double* dst = kv.data() + (size_t)k * b2;
    const bool tr = (m & 0x80000000u) != 0;
    for (int i = 0; i < nd; i++)
      for (int j = 0; j < nd; j++) dst[i * nd + j] = tr ? src[j * nd + i] : src[i * nd + j];
  }
  return 0;
}

int host_direct_solve(pm_fem* f, const std::vector<double>& b, std::vector<double>& x) {
  const int nd = f->ndim;
  const int64_t ndof = f->ndof;
  // K values to host (FP64, widened from FP32 for reduced-precision tangents)
  std::vector<double> kv;
  if (copy_kval_to_host(f, kv)) { f->err = "HIP error while copying K to the host"; return 1; }
  // free DOF numbering in RCM node order
  const int64_t np = f->npoint;
  std::vector<int32_t> order;
  order.reserve(np);
  {
    std::vector<char> seen(np, 0);
    auto deg = [&](int64_t a) { return f->h_browptr[a + 1] - f->h_browptr[a]; };
    while ((int64_t)order.size() < np) {
      int64_t start = -1;
      for (int64_t i = 0; i < np; i++) if (!seen[i] && (start < 0 || deg(i) < deg(start))) start = i;
      size_t head = order.size();
      order.push_back((int32_t)start); seen[start] = 1;
      while (head < order.size()) {
        int32_t v = order[head++];
        std::vector<int32_t> nb;
        for (int32_t k = f->h_browptr[v]; k < f->h_browptr[v + 1]; k++) {
          int32_t w = f->h_bcols[k];
          if (!seen[w]) { seen[w] = 1; nb.push_back(w); }
        }
        std::sort(nb.begin(), nb.end(), [&](int32_t p, int32_t q) { return deg(p) < deg(q) || (deg(p) == deg(q) && p < q); });
        for (auto w : nb) order.push_back(w);
      }
    }
    std::reverse(order.begin(), order.end());
  }
  std::vector<int64_t> pos(ndof, -1);
  std::vector<int64_t> eqs;
  for (auto a : order) for (int d = 0; d < nd; d++) {
    int64_t eq = (int64_t)nd * a + d;
    if (!f->is_prescribed[eq]) { pos[eq] = (int64_t)eqs.size(); eqs.push_back(eq); }
  }
  const int64_t n = (int64_t)eqs.size();
  // envelope: first column of each row
  std::vector<int64_t> first(n);
  for (int64_t i = 0; i < n; i++) first[i] = i;
  for (int64_t a = 0; a < np; a++)
    for (int32_t k = f->h_browptr[a]; k < f->h_browptr[a + 1]; k++) {
      int32_t bcol = f->h_bcols[k];
      for (int i = 0; i < nd; i++) for (int j = 0; j < nd; j++) {
        int64_t r = pos[(int64_t)nd * a + i], c = pos[(int64_t)nd * bcol + j];
        if (r < 0 || c < 0) continue;
        if (c < first[r]) first[r] = c;
      }
    }
  std::vector<int64_t> rptr(n + 1, 0);
  for (int64_t i = 0; i < n; i++) rptr[i + 1] = rptr[i] + (i - first[i] + 1);
  if (rptr[n] > (int64_t)4e9) { f->err = "host direct solver: envelope too large"; return 1; }
  std::vector<double> L(rptr[n], 0.0);
  auto at = [&](int64_t r, int64_t c) -> double& { return L[rptr[r] + (c - first[r])]; };
  for (int64_t a = 0; a < np; a++)
    for (int32_t k = f->h_browptr[a]; k < f->h_browptr[a + 1]; k++) {
      int32_t bcol = f->h_bcols[k];
      for (int i = 0; i < nd; i++) for (int j = 0; j < nd; j++) {
        int64_t r = pos[(int64_t)nd * a + i], c = pos[(int64_t)nd * bcol + j];
        if (r < 0 || c < 0 || c > r) continue;
        at(r, c) = kv[(size_t)k * nd * nd + i * nd + j];
      }
    }
  // Cholesky in envelope storage (row-oriented)
  for (int64_t i = 0; i < n; i++) {
    for (int64_t j = first[i]; j <= i; j++) {
      int64_t k0 = std::max(first[i], first[j]);
      double s = at(i, j);
      const double* li = &L[rptr[i] + (k0 - first[i])];
      const double* lj = &L[rptr[j] + (k0 - first[j])];
      for (int64_t k = 0; k < j - k0; k++) s -= li[k] * lj[k];
      if (j == i) {
        if (s <= 0.0) { f->err = "host direct solver: matrix is not positive definite"; return 1; }
        at(i, i) = std::sqrt(s);
      } else {
        at(i, j) = s / at(j, j);
      }
    }
  }
  auto cholesky_solve = [&](std::vector<double>& y) {  // y: right-hand side in, solution out
    for (int64_t i = 0; i < n; i++) {
      double s = y[i];
      for (int64_t k = first[i]; k < i; k++) s -= at(i, k) * y[k];
      y[i] = s / at(i, i);
    }
    for (int64_t i = n - 1; i >= 0; i--) {
      y[i] /= at(i, i);
      for (int64_t k = first[i]; k < i; k++) y[k] -= at(i, k) * y[i];
    }
  };
  std::vector<double> y(n);
  for (int64_t i = 0; i < n; i++) y[i] = b[eqs[i]];
  cholesky_solve(y);
  x.assign(ndof, 0.0);
  for (int64_t i = 0; i < n; i++) x[eqs[i]] = y[i];
  // two steps of iterative refinement against K_ff (as UMFPACK's default, UMFPACK_IRSTEP = 2, behind
  // pmsim's default direct solver; src/fem/linear_system.rs:185-201)
  std::vector<double> r(ndof);
  for (int step = 0; step < 2; step++) {
    for (int64_t a = 0; a < np; a++)
      for (int i = 0; i < nd; i++) {
        const int64_t eq = (int64_t)nd * a + i;
        if (pos[eq] < 0) continue;
        double s = b[eq];
        for (int32_t k = f->h_browptr[a]; k < f->h_browptr[a + 1]; k++) {
          const int64_t c0 = (int64_t)nd * f->h_bcols[k];
          for (int j = 0; j < nd; j++)
            if (pos[c0 + j] >= 0) s -= kv[(size_t)k * nd * nd + i * nd + j] * x[c0 + j];
        }
        r[eq] = s;
      }
    for (int64_t i = 0; i < n; i++) y[i] = r[eqs[i]];
    cholesky_solve(y);
    for (int64_t i = 0; i < n; i++) x[eqs[i]] += y[i];
  }
  return 0;
}

}  // namespace pm
