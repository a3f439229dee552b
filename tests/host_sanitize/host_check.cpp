// Host-side checks of the library's set-up code under AddressSanitizer and UndefinedBehaviorSanitizer (no GPU):
// mesh tables, element groups, node-block CSR patterns and gather maps, external forces, footprint estimates,
// integration rules, shape functions and extrapolation matrices, and the refusals of the C ABI.  Built from the
// library's own sources with the host code instrumented (test_host_sanitizers.py).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <set>
#include <vector>

#include "../../pmsim_amd/csrc/pm_internal.hpp"

namespace {

int failures = 0;
#define CHECK(c)                                                                   \
  do {                                                                             \
    if (!(c)) { std::printf("FAILED %s:%d: %s\n", __FILE__, __LINE__, #c); failures++; } \
  } while (0)

struct MeshData {
  int ndim = 2;
  std::vector<double> X;
  std::vector<int32_t> kinds, attrs, pts;
  std::vector<int64_t> ptr{0};
  int64_t npoint() const { return (int64_t)X.size() / ndim; }
  void cell(int kind, std::initializer_list<int32_t> p) {
    kinds.push_back(kind); attrs.push_back(1);
    pts.insert(pts.end(), p);
    ptr.push_back((int64_t)pts.size());
  }
};

// structured nx x ny Qua8 block of lx x ly: corner nodes on a (2nx+1) x (2ny+1) lattice without the cell centres
MeshData qua8_block(int nx, int ny, double lx, double ly) {
  MeshData m;
  std::map<std::pair<int, int>, int32_t> id;
  auto node = [&](int i, int j) {
    auto it = id.find({i, j});
    if (it != id.end()) return it->second;
    const int32_t k = (int32_t)(m.X.size() / 2);
    m.X.push_back(lx * i / (2.0 * nx)); m.X.push_back(ly * j / (2.0 * ny));
    id[{i, j}] = k;
    return k;
  };
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++) {
      const int a = 2 * i, b = 2 * j;
      m.cell(pm::QUA8, {node(a, b), node(a + 2, b), node(a + 2, b + 2), node(a, b + 2), node(a + 1, b), node(a + 2, b + 1),
                        node(a + 1, b + 2), node(a, b + 1)});
    }
  return m;
}

// nx x ny x nz Hex20 block of the unit cube (gemlab's local order: corners, bottom mid-edges, top mid-edges, verticals)
MeshData hex20_block(int nx, int ny, int nz) {
  MeshData m;
  m.ndim = 3;
  std::map<std::tuple<int, int, int>, int32_t> id;
  auto node = [&](int i, int j, int k) {
    auto it = id.find({i, j, k});
    if (it != id.end()) return it->second;
    const int32_t n = (int32_t)(m.X.size() / 3);
    m.X.push_back(i / (2.0 * nx)); m.X.push_back(j / (2.0 * ny)); m.X.push_back(k / (2.0 * nz));
    id[{i, j, k}] = n;
    return n;
  };
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        const int a = 2 * i, b = 2 * j, c = 2 * k;
        m.cell(pm::HEX20, {node(a, b, c), node(a + 2, b, c), node(a + 2, b + 2, c), node(a, b + 2, c),
                           node(a, b, c + 2), node(a + 2, b, c + 2), node(a + 2, b + 2, c + 2), node(a, b + 2, c + 2),
                           node(a + 1, b, c), node(a + 2, b + 1, c), node(a + 1, b + 2, c), node(a, b + 1, c),
                           node(a + 1, b, c + 2), node(a + 2, b + 1, c + 2), node(a + 1, b + 2, c + 2), node(a, b + 1, c + 2),
                           node(a, b, c + 1), node(a + 2, b, c + 1), node(a + 2, b + 2, c + 1), node(a, b + 2, c + 1)});
      }
  return m;
}

// a 2D mesh mixing Tri3 pairs and Qua4 cells on an n x n lattice
MeshData mixed_tri3_qua4(int n) {
  MeshData m;
  for (int j = 0; j <= n; j++)
    for (int i = 0; i <= n; i++) { m.X.push_back(i); m.X.push_back(j); }
  auto v = [&](int i, int j) { return (int32_t)(j * (n + 1) + i); };
  for (int j = 0; j < n; j++)
    for (int i = 0; i < n; i++) {
      if ((i + j) % 2) { m.cell(pm::TRI3, {v(i, j), v(i + 1, j), v(i + 1, j + 1)}); m.cell(pm::TRI3, {v(i, j), v(i + 1, j + 1), v(i, j + 1)}); }
      else m.cell(pm::QUA4, {v(i, j), v(i + 1, j), v(i + 1, j + 1), v(i, j + 1)});
    }
  return m;
}

pm_fem* create(const MeshData& m) {
  pm_fem* f = nullptr;
  const int rc = pm_fem_create(&f, m.ndim, m.npoint(), m.X.data(), (int64_t)m.kinds.size(), m.kinds.data(), m.attrs.data(),
                               m.ptr.data(), m.pts.data());
  CHECK(rc == 0 && f);
  return f;
}

// pattern and map invariants after the host set-up
void check_tables(pm_fem* f, double expect_fz_total, int load_dof) {
  CHECK(pm::build_groups(f) == 0);
  CHECK(pm::build_patterns(f) == 0);
  const int64_t np = f->npoint;
  CHECK((int64_t)f->h_browptr.size() == np + 1);
  int64_t diag = 0;
  std::set<std::pair<int32_t, int32_t>> blocks;
  for (int64_t a = 0; a < np; a++) {
    for (int32_t k = f->h_browptr[a]; k < f->h_browptr[a + 1]; k++) {
      if (k > f->h_browptr[a]) CHECK(f->h_bcols[k] > f->h_bcols[k - 1]);  // sorted, unique
      CHECK(f->h_bcols[k] >= 0 && f->h_bcols[k] < np);
      blocks.insert({(int32_t)a, f->h_bcols[k]});
      if (f->h_bcols[k] == a) diag++;
    }
  }
  CHECK(diag == np);
  for (const auto& b : blocks) CHECK(blocks.count({b.second, b.first}) == 1);  // symmetric pattern
  CHECK(f->nnzu == (f->nnzb + np) / 2);
  CHECK((int64_t)f->h_kmap_ptr.size() == f->nnzu + 1);
  for (int64_t i = 0; i < f->nnzu; i++) CHECK(f->h_kmap_ptr[i + 1] > f->h_kmap_ptr[i]);  // every stored block has a contribution
  for (auto v : f->h_kmap) CHECK((v & 0x7fffffff) < f->n_ke_values);
  CHECK((int64_t)f->h_fmap_ptr.size() == np + 1);
  for (auto v : f->h_fmap) CHECK((int64_t)v * f->ndim < f->n_fe);
  pm::assemble_f_ext(f, 0);
  double s = 0.0;
  for (int64_t a = 0; a < np; a++) s += f->f_ext[(size_t)a * f->ndim + load_dof];
  CHECK(std::fabs(s - expect_fz_total) < 1e-10 * std::fabs(expect_fz_total));
  int64_t bytes = 0;
  CHECK(pm_estimate_device_bytes(f, &bytes) == 0 && bytes > 0);
}

}  // namespace

int main() {
  // 2D footing-like block: von Mises, bottom fixed, Qy on the top edges (total = q * width)
  {
    MeshData m = qua8_block(12, 6, 10.0, 5.0);
    pm_fem* f = create(m);
    CHECK(pm_set_param_solid(f, 1, 1, 1e4, 0.3, 0.0, 200.0, 9) == 0);
    std::vector<int32_t> ep, ed;
    std::vector<double> tab;
    for (int64_t p = 0; p < m.npoint(); p++)
      if (m.X[2 * p + 1] == 0.0) for (int d = 0; d < 2; d++) { ep.push_back((int32_t)p); ed.push_back(d); tab.push_back(0.0); }
    CHECK(pm_set_essential(f, (int64_t)ep.size(), ep.data(), ed.data(), tab.data(), 1) == 0);
    for (int64_t e = 0; e < (int64_t)m.kinds.size(); e++) {
      const int32_t* c = &m.pts[m.ptr[e]];
      if (m.X[2 * c[2] + 1] == 5.0) {
        const int32_t edge[3] = {c[3], c[2], c[6]};
        const double q[2] = {-3.0, -6.0};
        CHECK(pm_add_facet_load(f, 2, pm::LIN3, 3, edge, e, q, 2) == 0);
      }
    }
    check_tables(f, -3.0 * 10.0, 1);
    pm_fem_destroy(f);
  }
  // 3D Hex20 block: Qz on the top faces (total = q), Qn refused on a 3D edge, Qz refused in 2D below
  {
    MeshData m = hex20_block(4, 3, 2);
    pm_fem* f = create(m);
    CHECK(pm_set_param_solid(f, 1, 1, 2e5, 0.3, 1e3, 250.0, 27) == 0);
    for (int64_t e = 0; e < (int64_t)m.kinds.size(); e++) {
      const int32_t* c = &m.pts[m.ptr[e]];
      if (m.X[3 * c[4] + 2] == 1.0) {
        const int32_t face[8] = {c[4], c[5], c[6], c[7], c[12], c[13], c[14], c[15]};
        const double q = 2.5;
        CHECK(pm_add_facet_load(f, 3, pm::QUA8, 8, face, e, &q, 1) == 0);
      }
    }
    const int32_t edge[2] = {0, 1};
    const double q = 1.0;
    CHECK(pm_add_facet_load(f, 0, pm::LIN2, 2, edge, 0, &q, 1) != 0);
    check_tables(f, 2.5, 2);
    pm_fem_destroy(f);
  }
  // mixed Tri3 / Qua4 mesh, linear elastic, point loads
  {
    MeshData m = mixed_tri3_qua4(7);
    pm_fem* f = create(m);
    CHECK(pm_set_param_solid(f, 1, 0, 1000.0, 0.25, 0.0, 0.0, 0) == 0);
    const int32_t pp[2] = {3, 5}, pd[2] = {1, 1};
    const double pv[2] = {-1.0, -2.0};
    CHECK(pm_set_point_loads(f, 2, pp, pd, pv, 1) == 0);
    const int32_t edge[2] = {0, 1};
    CHECK(pm_add_facet_load(f, 3, pm::LIN2, 2, edge, 0, pv, 1) != 0);  // no Uz in 2D
    check_tables(f, -3.0, 1);
    pm_fem_destroy(f);
  }
  // integration rules, shape functions, extrapolation matrices of every kind and rule
  const int kinds[] = {pm::TRI3, pm::TRI6, pm::TRI10, pm::TRI15, pm::QUA4, pm::QUA8, pm::QUA9, pm::TET4, pm::TET10, pm::HEX8, pm::HEX20};
  for (int k : kinds)
    for (int ng : {0, 1, 3, 4, 6, 7, 8, 9, 12, 16, 27}) {
      int32_t n = 0;
      if (pm_gauss_rule(k, ng, &n, nullptr, nullptr) != 0) continue;
      std::vector<double> xi((size_t)n * 3), w(n), N((size_t)n * pm::kind_nnode(k));
      CHECK(pm_gauss_rule(k, ng, &n, xi.data(), w.data()) == 0);
      CHECK(pm_shape_values(k, n, xi.data(), N.data()) == 0);
      for (int q = 0; q < n; q++) {
        double s = 0.0;
        for (int a = 0; a < pm::kind_nnode(k); a++) s += N[(size_t)q * pm::kind_nnode(k) + a];
        CHECK(std::fabs(s - 1.0) < 1e-13);  // partition of unity
      }
      std::vector<double> E((size_t)pm::kind_nnode(k) * n);
      (void)pm_extrapolation_matrix(k, ng, E.data());
    }
  // refusals without a device
  pm_fem* f = create(qua8_block(2, 1, 1.0, 1.0));
  CHECK(pm_initialize(f, 0) != 0);
  CHECK(pm_set_param_solid(f, 1, 7, 1.0, 0.3, 0.0, 1.0, 0) != 0);
  pm_fem_destroy(f);
  std::printf("host checks: %d failures\n", failures);
  return failures ? 1 : 0;
}
