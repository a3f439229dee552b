// ORACLE — test infrastructure only.
//
// Serial FP64 CPU restatement of pmsim's implicit solid element loop, used as the
// parity checker for the HIP path and as the timed CPU baseline.  Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load this library;
// the product library never links or calls it.
//
// What it restates (pmsim 0.3.0, paths relative to the reference root):
//   * DOF numbering eq = ndim*p + d          src/base/all_dofs.rs:87-117, element_dofs.rs:123-131
//   * l2g maps and assembly with `ignore`      src/base/assembly.rs:8-18, 38-46, 70-139
//   * strain increment at an IP                src/base/calculate_strain.rs:18-60
//   * f_int = sum B^T sigma dV, K = B^T D B dV src/fem/element_solid.rs:124-138, 168-185
//   * update/backup/restore/reset              src/fem/element_solid.rs:190-246, material/local_state.rs:40-44
//   * von Mises return map + tangent           src/material/von_mises.rs:107-226
//   * linear elasticity                        src/material/linear_elastic.rs:49-71
//   * prescribed values (LMM / RSM)            src/fem/bc_prescribed.rs:61-157
//   * point loads, normal distributed loads    src/fem/bc_concentrated.rs:46-55, bc_distributed.rs:166-174
//   * residual, primary update                 src/fem/solver_data.rs:85-152
//   * Newton driver and time loop              src/fem/solver_implicit.rs:56-272, time_control.rs:66-90
//   * convergence tests                        src/fem/convergence_control.rs:52-127
//
// Independent constructions (deliberately different from the HIP product):
//   * shape functions come from inverting the Vandermonde matrix of each element's
//     monomial basis at its reference nodes (the product uses closed forms);
//   * the global system is solved LITERALLY as pmsim builds it (augmented LMM
//     matrix or RSM matrix with unit diagonal) by a banded LU with partial
//     pivoting after a reverse Cuthill-McKee node ordering (the product condenses
//     the prescribed DOFs and runs PCG or a skyline Cholesky);
//   * element tangents are the dense triple product B^T (D B).
//
// Parity pinning of each part is documented in the tests that load this file.
// Parity unpinned: Tri10 node ordering and the default Gauss rules of Tri6/Tri10/
// Tet10 (gemlab conventions absent from the reference; no fixture covers them).

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr double SQ2 = 1.41421356237309504880;

// ---------------------------------------------------------------------------------------------
// element kinds: reference nodes + monomial basis  (gemlab GeoKind conventions)
// ---------------------------------------------------------------------------------------------
enum Kind { TRI3 = 0, TRI6, TRI10, TRI15, QUA4, QUA8, QUA9, TET4, TET10, HEX8, HEX20, NKIND,
            // facet kinds used only for boundary loads
            LIN2 = 100, LIN3, LIN4, LIN5 };

struct Basis {
  int ndim = 0, nnode = 0, family = 0;  // family 0 tri, 1 qua, 2 tet, 3 hex, 4 lin
  std::vector<std::array<double, 3>> nodes;
  std::vector<std::array<int, 3>> mono;
  std::vector<double> C;  // N_i(x) = sum_k C[k*n+i] * mono_k(x)
};

double ipow(double x, int e) { double r = 1; for (int i = 0; i < e; i++) r *= x; return r; }

// Solve A X = B (dense, partial pivoting); A n×n, B n×m, both row-major; result in B.
bool dense_solve(std::vector<double> A, std::vector<double>& B, int n, int m) {
  for (int k = 0; k < n; k++) {
    int piv = k; double best = std::fabs(A[k * n + k]);
    for (int i = k + 1; i < n; i++) if (std::fabs(A[i * n + k]) > best) { best = std::fabs(A[i * n + k]); piv = i; }
    if (best == 0.0) return false;
    if (piv != k) {
      for (int j = 0; j < n; j++) std::swap(A[k * n + j], A[piv * n + j]);
      for (int j = 0; j < m; j++) std::swap(B[k * m + j], B[piv * m + j]);
    }
    for (int i = k + 1; i < n; i++) {
      double f = A[i * n + k] / A[k * n + k];
      if (f == 0.0) continue;
      for (int j = k; j < n; j++) A[i * n + j] -= f * A[k * n + j];
      for (int j = 0; j < m; j++) B[i * m + j] -= f * B[k * m + j];
    }
  }
  for (int k = n - 1; k >= 0; k--) {
    for (int j = 0; j < m; j++) {
      double s = B[k * m + j];
      for (int i = k + 1; i < n; i++) s -= A[k * n + i] * B[i * m + j];
      B[k * m + j] = s / A[k * n + k];
    }
  }
  return true;
}

Basis make_basis(int kind) {
  Basis b;
  auto tri_mono = [&](int p) { for (int d = 0; d <= p; d++) for (int a = d; a >= 0; a--) b.mono.push_back({a, d - a, 0}); };
  auto tet_mono = [&](int p) {
    for (int a = 0; a <= p; a++) for (int c = 0; c <= p - a; c++) for (int e = 0; e <= p - a - c; e++) b.mono.push_back({a, c, e});
  };
  switch (kind) {
    case TRI3: b.ndim = 2; b.family = 0; b.nodes = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}}; tri_mono(1); break;
    case TRI6: b.ndim = 2; b.family = 0;
      b.nodes = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {.5, 0, 0}, {.5, .5, 0}, {0, .5, 0}}; tri_mono(2); break;
    case TRI10: b.ndim = 2; b.family = 0;
      b.nodes = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {1. / 3, 0, 0}, {2. / 3, 1. / 3, 0}, {0, 2. / 3, 0},
                 {2. / 3, 0, 0}, {1. / 3, 2. / 3, 0}, {0, 1. / 3, 0}, {1. / 3, 1. / 3, 0}};
      tri_mono(3); break;
    case TRI15: b.ndim = 2; b.family = 0;
      b.nodes = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {.5, 0, 0}, {.5, .5, 0}, {0, .5, 0}, {.25, 0, 0}, {.75, 0, 0},
                 {.75, .25, 0}, {.25, .75, 0}, {0, .75, 0}, {0, .25, 0}, {.25, .25, 0}, {.5, .25, 0}, {.25, .5, 0}};
      tri_mono(4); break;
    case QUA4: case QUA8: case QUA9:
      b.ndim = 2; b.family = 1;
      b.nodes = {{-1, -1, 0}, {1, -1, 0}, {1, 1, 0}, {-1, 1, 0}};
      if (kind != QUA4) { b.nodes.push_back({0, -1, 0}); b.nodes.push_back({1, 0, 0}); b.nodes.push_back({0, 1, 0}); b.nodes.push_back({-1, 0, 0}); }
      if (kind == QUA9) b.nodes.push_back({0, 0, 0});
      if (kind == QUA4) b.mono = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {1, 1, 0}};
      else if (kind == QUA8) b.mono = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {2, 0, 0}, {1, 1, 0}, {0, 2, 0}, {2, 1, 0}, {1, 2, 0}};
      else for (int j = 0; j < 3; j++) for (int i = 0; i < 3; i++) b.mono.push_back({i, j, 0});
      break;
    case TET4: b.ndim = 3; b.family = 2; b.nodes = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}}; tet_mono(1); break;
    case TET10: b.ndim = 3; b.family = 2;
      b.nodes = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {.5, 0, 0}, {.5, .5, 0}, {0, .5, 0}, {0, 0, .5}, {.5, 0, .5}, {0, .5, .5}};
      tet_mono(2); break;
    case HEX8: case HEX20:
      b.ndim = 3; b.family = 3;
      b.nodes = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1}, {-1, -1, 1}, {1, -1, 1}, {1, 1, 1}, {-1, 1, 1}};
      if (kind == HEX20) {
        std::vector<std::array<double, 3>> mid = {{0, -1, -1}, {1, 0, -1}, {0, 1, -1}, {-1, 0, -1}, {0, -1, 1}, {1, 0, 1},
                                                  {0, 1, 1}, {-1, 0, 1}, {-1, -1, 0}, {1, -1, 0}, {1, 1, 0}, {-1, 1, 0}};
        for (auto& m : mid) b.nodes.push_back(m);
        // 20-term serendipity space
        for (int a = 0; a < 3; a++) for (int c = 0; c < 3; c++) for (int e = 0; e < 3; e++) {
          int n2 = (a == 2) + (c == 2) + (e == 2);
          int n1 = (a == 1) + (c == 1) + (e == 1);
          // keep monomials whose "superlinear" degree is at most 2
          int sdeg = (a == 2 ? 2 : 0) + (c == 2 ? 2 : 0) + (e == 2 ? 2 : 0);
          (void)n1;
          if (n2 == 0 || (n2 == 1 && sdeg == 2)) {
            // r^2 s^2 t^x style terms excluded; r^2 * (linear in others) allowed
            b.mono.push_back({a, c, e});
          }
        }
        // the loop above admits: all {0,1}^3 (8) + one squared variable times {0,1}^2 of the others (12) = 20
      } else {
        for (int e = 0; e < 2; e++) for (int c = 0; c < 2; c++) for (int a = 0; a < 2; a++) b.mono.push_back({a, c, e});
      }
      break;
    case LIN2: case LIN3: case LIN4: case LIN5: {
      b.ndim = 1; b.family = 4;
      int n = kind - LIN2 + 2;
      b.nodes.push_back({-1, 0, 0}); b.nodes.push_back({1, 0, 0});
      for (int i = 1; i < n - 1; i++) b.nodes.push_back({-1 + 2.0 * i / (n - 1), 0, 0});
      for (int a = 0; a < n; a++) b.mono.push_back({a, 0, 0});
      break;
    }
    default: break;
  }
  b.nnode = (int)b.nodes.size();
  int n = b.nnode;
  if ((int)b.mono.size() != n) { std::fprintf(stderr, "oracle: basis size mismatch kind %d\n", kind); std::abort(); }
  // V[i][k] = mono_k(node_i); N = V^{-T} m  =>  C = V^{-1}
  std::vector<double> V(n * n), I(n * n, 0.0);
  for (int i = 0; i < n; i++) {
    for (int k = 0; k < n; k++) {
      auto& e = b.mono[k]; auto& x = b.nodes[i];
      V[i * n + k] = ipow(x[0], e[0]) * ipow(x[1], e[1]) * ipow(x[2], e[2]);
    }
    I[i * n + i] = 1.0;
  }
  if (!dense_solve(V, I, n, n)) { std::fprintf(stderr, "oracle: singular Vandermonde kind %d\n", kind); std::abort(); }
  b.C = I;  // C[k*n+i]: coefficient of mono k in N_i
  return b;
}

const Basis& basis(int kind) {
  static std::map<int, Basis> cache;
  auto it = cache.find(kind);
  if (it == cache.end()) it = cache.emplace(kind, make_basis(kind)).first;
  return it->second;
}

// N[i], dN[i*ndim+j] at reference point x
void eval_basis(const Basis& b, const double* x, double* N, double* dN) {
  int n = b.nnode, nd = b.ndim;
  double m[32], dm[32 * 3];
  for (int k = 0; k < n; k++) {
    auto& e = b.mono[k];
    double p0 = ipow(x[0], e[0]), p1 = nd > 1 ? ipow(x[1], e[1]) : 1.0, p2 = nd > 2 ? ipow(x[2], e[2]) : 1.0;
    m[k] = p0 * p1 * p2;
    dm[k * 3 + 0] = e[0] ? e[0] * ipow(x[0], e[0] - 1) * p1 * p2 : 0.0;
    dm[k * 3 + 1] = (nd > 1 && e[1]) ? e[1] * ipow(x[1], e[1] - 1) * p0 * p2 : 0.0;
    dm[k * 3 + 2] = (nd > 2 && e[2]) ? e[2] * ipow(x[2], e[2] - 1) * p0 * p1 : 0.0;
  }
  for (int i = 0; i < n; i++) {
    double s = 0, d0 = 0, d1 = 0, d2 = 0;
    for (int k = 0; k < n; k++) {
      double c = b.C[k * n + i];
      s += c * m[k]; d0 += c * dm[k * 3]; d1 += c * dm[k * 3 + 1]; d2 += c * dm[k * 3 + 2];
    }
    N[i] = s;
    dN[i * nd + 0] = d0;
    if (nd > 1) dN[i * nd + 1] = d1;
    if (nd > 2) dN[i * nd + 2] = d2;
  }
}

// ---------------------------------------------------------------------------------------------
// Gauss rules (gemlab integ::Gauss conventions; r fastest for tensor rules)
// ---------------------------------------------------------------------------------------------
struct IP { double x[3]; double w; };

std::vector<IP> gauss_legendre_1d(int n) {
  // Newton iteration on Legendre polynomials (computed, not tabulated)
  std::vector<IP> r(n);
  for (int i = 0; i < n; i++) {
    double x = std::cos(M_PI * (i + 0.75) / (n + 0.5));
    for (int it = 0; it < 100; it++) {
      double p0 = 1, p1 = x;
      for (int k = 2; k <= n; k++) { double p2 = ((2 * k - 1) * x * p1 - (k - 1) * p0) / k; p0 = p1; p1 = p2; }
      double pn = (n == 1) ? x : p1, pnm1 = (n == 1) ? 1.0 : p0;
      double dp = n * (x * pn - pnm1) / (x * x - 1);
      double dx = pn / dp; x -= dx;
      if (std::fabs(dx) < 1e-17) break;
    }
    double p0 = 1, p1 = x;
    for (int k = 2; k <= n; k++) { double p2 = ((2 * k - 1) * x * p1 - (k - 1) * p0) / k; p0 = p1; p1 = p2; }
    double pn = (n == 1) ? x : p1, pnm1 = (n == 1) ? 1.0 : p0;
    double dp = n * (x * pn - pnm1) / (x * x - 1);
    r[n - 1 - i] = {{x, 0, 0}, 2.0 / ((1 - x * x) * dp * dp)};
  }
  return r;
}

std::vector<IP> tri_rule(int n) {
  std::vector<IP> r;
  auto add3 = [&](double a, double b, double w) {  // permutations of barycentric (a, a, b) -> (r, s)
    r.push_back({{a, a, 0}, w}); r.push_back({{b, a, 0}, w}); r.push_back({{a, b, 0}, w});
  };
  auto add6 = [&](double a, double b, double c, double w) {
    r.push_back({{a, b, 0}, w}); r.push_back({{b, a, 0}, w}); r.push_back({{b, c, 0}, w});
    r.push_back({{c, b, 0}, w}); r.push_back({{c, a, 0}, w}); r.push_back({{a, c, 0}, w});
  };
  if (n == 1) r.push_back({{1. / 3, 1. / 3, 0}, 0.5});
  else if (n == 3) { r.push_back({{1. / 6, 1. / 6, 0}, 1. / 6}); r.push_back({{2. / 3, 1. / 6, 0}, 1. / 6}); r.push_back({{1. / 6, 2. / 3, 0}, 1. / 6}); }
  else if (n == 6) {  // degree 4 (Strang-Fix / Dunavant)
    add3(0.445948490915965, 0.10810301816807, 0.5 * 0.223381589678011);
    add3(0.091576213509771, 0.816847572980459, 0.5 * 0.109951743655322);
  } else if (n == 7) {  // degree 5
    r.push_back({{1. / 3, 1. / 3, 0}, 0.5 * 0.225});
    add3(0.470142064105115, 0.059715871789770, 0.5 * 0.132394152788506);
    add3(0.101286507323456, 0.797426985353087, 0.5 * 0.125939180544827);
  } else if (n == 12) {  // degree 6 (Dunavant)
    add3(0.249286745170910, 0.501426509658179, 0.5 * 0.116786275726379);
    add3(0.063089014491502, 0.873821971016996, 0.5 * 0.050844906370207);
    add6(0.053145049844817, 0.310352451033784, 0.636502499121399, 0.5 * 0.082851075618374);
  }
  return r;
}

std::vector<IP> tet_rule(int n) {
  std::vector<IP> r;
  if (n == 1) r.push_back({{0.25, 0.25, 0.25}, 1. / 6});
  else if (n == 4) {
    double a = (5.0 - std::sqrt(5.0)) / 20.0, b = (5.0 + 3.0 * std::sqrt(5.0)) / 20.0;
    r.push_back({{a, a, a}, 1. / 24}); r.push_back({{b, a, a}, 1. / 24}); r.push_back({{a, b, a}, 1. / 24}); r.push_back({{a, a, b}, 1. / 24});
  } else if (n == 5) {
    r.push_back({{.25, .25, .25}, -2. / 15});
    double a = 1. / 6, b = 0.5, w = 3. / 40;
    r.push_back({{a, a, a}, w}); r.push_back({{b, a, a}, w}); r.push_back({{a, b, a}, w}); r.push_back({{a, a, b}, w});
  }
  return r;
}

std::vector<IP> gauss_rule(int kind, int ngauss) {
  const Basis& b = basis(kind);
  std::vector<IP> r;
  if (b.family == 1 || b.family == 3 || b.family == 4) {
    int nd = b.ndim, n1 = 0;
    for (int k = 1; k <= 6; k++) { int t = 1; for (int d = 0; d < nd; d++) t *= k; if (t == ngauss) n1 = k; }
    if (n1 == 0) return r;
    auto g = gauss_legendre_1d(n1);
    if (nd == 1) return g;
    if (nd == 2) { for (int j = 0; j < n1; j++) for (int i = 0; i < n1; i++) r.push_back({{g[i].x[0], g[j].x[0], 0}, g[i].w * g[j].w}); }
    else for (int k = 0; k < n1; k++) for (int j = 0; j < n1; j++) for (int i = 0; i < n1; i++)
      r.push_back({{g[i].x[0], g[j].x[0], g[k].x[0]}, g[i].w * g[j].w * g[k].w});
    return r;
  }
  if (b.family == 0) return tri_rule(ngauss);
  return tet_rule(ngauss);
}

// default rules when ParamSolid.ngauss is None; Tri3's 3 points are what pmsim writes in its own state
// files (data/results/artificial/artificial-elastic-2d-*.json: "ngauss": 3 for every Tri3 cell of a
// base whose ngauss is null), Qua8 9 and Hex8 8 likewise (artificial-elastic-2d-qua8, -3d)
int default_ngauss(int kind) {
  switch (kind) {
    case TRI3: return 3; case TRI6: return 3; case TRI10: return 6; case TRI15: return 12;
    case QUA4: return 4; case QUA8: return 9; case QUA9: return 9;
    case TET4: return 1; case TET10: return 4; case HEX8: return 8; case HEX20: return 27;
  }
  return 0;
}

// ---------------------------------------------------------------------------------------------
// materials (Mandel basis)
// ---------------------------------------------------------------------------------------------
struct Param { int model = 0; double E = 0, nu = 0, H = 0, z0 = 0; int ngauss = 0; double K = 0, G = 0; };

struct Ideal { int ndim = 2; bool axisym = false, plane_stress = false; double thickness = 1.0; };

// elastic modulus matrix D (nsig×nsig); plane stress per LinearElastic only
void elastic_D(const Param& p, const Ideal& id, int ns, double* D) {
  for (int i = 0; i < ns * ns; i++) D[i] = 0;
  if (id.ndim == 2 && id.plane_stress) {
    double c = p.E / (1 - p.nu * p.nu);
    D[0] = c; D[1] = c * p.nu; D[ns] = c * p.nu; D[ns + 1] = c; D[3 * ns + 3] = c * (1 - p.nu);
    return;
  }
  for (int i = 0; i < ns; i++) for (int j = 0; j < ns; j++) {
    double Ii = i < 3 ? 1.0 : 0.0, Ij = j < 3 ? 1.0 : 0.0;
    double psd = (i == j ? 1.0 : 0.0) - Ii * Ij / 3.0;
    D[i * ns + j] = 2.0 * p.G * psd + p.K * Ii * Ij;
  }
}

double sigma_m(const double* s) { return (s[0] + s[1] + s[2]) / 3.0; }
double sigma_d(const double* s, int ns) {
  double m = sigma_m(s), ss = 0;
  for (int i = 0; i < ns; i++) { double d = s[i] - (i < 3 ? m : 0.0); ss += d * d; }
  return std::sqrt(1.5 * ss);
}

// IP state: sigma[ns], z, lambda, elastic, epbar
struct IPState { double sig[6] = {0, 0, 0, 0, 0, 0}; double z = 0, lam = 0; bool elastic = true; double epbar = 0; };

const char* ERR_VM_ZERO = "von Mises plastic update must not lead to zero sigma_d (sigma_d < 1e-10)";
const char* ERR_VM_STIFF = "von Mises stiffness cannot be computed with zero deviatoric norm (norm_s < 1e-10)";

// von_mises.rs:170-226 / linear_elastic.rs update
const char* update_stress(const Param& p, const Ideal& id, int ns, IPState& st, const double* deps) {
  double D[36];
  elastic_D(p, id, ns, D);
  if (p.model == 0) {
    for (int i = 0; i < ns; i++) { double s = 0; for (int j = 0; j < ns; j++) s += D[i * ns + j] * deps[j]; st.sig[i] += s; }
    return nullptr;
  }
  st.elastic = true;
  st.lam = 0.0;
  double tr[6];
  for (int i = 0; i < ns; i++) { double s = 0; for (int j = 0; j < ns; j++) s += D[i * ns + j] * deps[j]; tr[i] = st.sig[i] + s; }
  for (int i = 0; i < ns; i++) st.sig[i] = tr[i];
  double ftr = sigma_d(st.sig, ns) - st.z;
  if (ftr / p.z0 <= 1e-6) return nullptr;
  double sm = sigma_m(st.sig), sd = sigma_d(st.sig, ns);
  double lam = ftr / (3.0 * p.G + p.H);
  double beta = 1.0 - lam * 3.0 * p.G / sd;
  for (int i = 0; i < ns; i++) { double iv = i < 3 ? 1.0 : 0.0; double dev = st.sig[i] - sm * iv; st.sig[i] = sm * iv + beta * dev; }
  if (sigma_d(st.sig, ns) < 1e-10) return ERR_VM_ZERO;
  st.z += p.H * lam;
  st.elastic = false;
  st.lam = lam;
  return nullptr;
}

// von_mises.rs:124-167
const char* stiffness(const Param& p, const Ideal& id, int ns, const IPState& st, double* D) {
  if (p.model == 0 || st.elastic) { elastic_D(p, id, ns, D); return nullptr; }
  double sm = sigma_m(st.sig), sd = sigma_d(st.sig, ns);
  double s[6];
  for (int i = 0; i < ns; i++) s[i] = st.sig[i] - (i < 3 ? sm : 0.0);
  double sdtr = sd + st.lam * 3.0 * p.G;
  double norm_s = sd * std::sqrt(2.0 / 3.0);
  if (norm_s < 1e-10) return ERR_VM_STIFF;
  double d = 3.0 * p.G + p.H;
  double a = 2.0 * p.G * (1.0 - st.lam * 3.0 * p.G / sdtr);
  double bb = 6.0 * p.G * p.G * (st.lam / sdtr - 1.0 / d) / (norm_s * norm_s);
  for (int i = 0; i < ns; i++) for (int j = 0; j < ns; j++) {
    double Ii = i < 3 ? 1.0 : 0.0, Ij = j < 3 ? 1.0 : 0.0;
    double psd = (i == j ? 1.0 : 0.0) - Ii * Ij / 3.0;
    D[i * ns + j] = a * psd + bb * s[i] * s[j] + p.K * Ii * Ij;
  }
  return nullptr;
}

// ---------------------------------------------------------------------------------------------
// problem
// ---------------------------------------------------------------------------------------------
struct Facet { int kind; std::vector<int> pts; int cell; std::vector<double> values; int nbc = 0; };  // nbc 0 Qn, 1-3 Qx/Qy/Qz

struct Problem {
  int ndim = 2, npoint = 0, ncell = 0;
  std::vector<double> X;
  std::vector<int> kind, attr, cptr, cpts;
  std::map<int, Param> params;
  Ideal ideal;
  // config (config.rs:184-231 defaults)
  bool lmm = false, linear_problem = false;
  bool arc = false;                   // arc-length method (config.rs:41-54; defaults ψ = 1, ℓ₀ = 1)
  double arc_psi = 1.0, ell0 = 1.0;
  double ell = 0.0;                   // loading factor ℓ (FemState::ell, 0 before the first step)
  double t_ini = 0, t_fin = 1, dt = 1, dt_out = 1, tol_rr_abs = 1e-10, tol_mdu_rel = 1e-8;
  int n_max_iterations = 10, n_max_time_steps = 1000, allowed_step_n_failure = 100;
  // prescribed: value(t_k) tables, k = 0..nstation-1 with t_k = t_ini + k*dt
  std::vector<int> ess_point, ess_dof;
  std::vector<double> ess_table;
  int nstation = 0;
  std::vector<int> pl_point, pl_dof; std::vector<double> pl_table;  // concentrated loads (n x pl_nstation)
  int pl_nstation = 1;
  std::vector<Facet> facets;                                         // Qn loads (values per station)
  // derived
  int ns = 4, ndof = 0, neq_total = 0, nip_total = 0;
  std::vector<int> ip_ptr;            // per-cell first IP index
  std::vector<std::vector<IP>> rules; // per cell (shared by kind/ngauss but kept simple)
  std::vector<IPState> st, backup;
  std::vector<double> u, ddu;
  std::vector<int> prescribed_eq;
  std::vector<char> ignore;
  double t = 0;
  std::string err;
  bool ready = false;
  // last-solve statistics
  std::vector<int> iters;
  std::vector<double> norm_rr_hist;
};

// value of a natural load at time t from its station values (one value: constant), the station
// being t_k = t_ini + k*dt; pmsim evaluates the load's function at the step's time
// (src/fem/bc_distributed.rs:161-164, src/fem/bc_concentrated.rs:46-55)
double load_value(const Problem& P, const double* v, int nst, double t) {
  int ik = (int)std::llround((t - P.t_ini) / P.dt);
  if (ik < 0) ik = 0;
  if (ik >= nst) ik = nst - 1;
  return v[ik];
}

double table_value(const Problem& P, int p, double t) {
  double k = (t - P.t_ini) / P.dt;
  int ik = (int)std::llround(k);
  if (ik < 0) ik = 0;
  if (ik >= P.nstation) ik = P.nstation - 1;
  return P.ess_table[(size_t)p * P.nstation + ik];
}

const char* setup(Problem& P) {
  if (P.ready) return nullptr;
  P.ns = P.ndim == 2 ? 4 : 6;
  P.ndof = P.ndim * P.npoint;
  P.neq_total = P.ndof + (P.lmm ? (int)P.ess_point.size() : 0);
  P.ip_ptr.assign(P.ncell + 1, 0);
  P.rules.resize(P.ncell);
  for (int e = 0; e < P.ncell; e++) {
    auto it = P.params.find(P.attr[e]);
    if (it == P.params.end()) return "cannot find attribute";
    Param& pa = it->second;
    int ng = pa.ngauss > 0 ? pa.ngauss : default_ngauss(P.kind[e]);
    P.rules[e] = gauss_rule(P.kind[e], ng);
    if (P.rules[e].empty()) return "requested number of integration points is not available";
    P.ip_ptr[e + 1] = P.ip_ptr[e] + (int)P.rules[e].size();
  }
  for (auto& kv : P.params) {
    Param& pa = kv.second;
    pa.K = pa.E / (3.0 * (1.0 - 2.0 * pa.nu));
    pa.G = pa.E / (2.0 * (1.0 + pa.nu));
    if (pa.model == 1) {
      if (P.ideal.plane_stress) return "von Mises model does not work in plane-stress";
      if (pa.z0 <= 1e-6) return "von Mises initial size of the yield surface must > 1e-6";
    }
  }
  P.nip_total = P.ip_ptr[P.ncell];
  P.st.assign(P.nip_total, IPState());
  P.backup = P.st;
  P.u.assign(P.neq_total, 0.0);
  P.ddu.assign(P.neq_total, 0.0);
  P.prescribed_eq.clear();
  P.ignore.assign(P.ndof, 0);
  for (size_t i = 0; i < P.ess_point.size(); i++) {
    int eq = P.ndim * P.ess_point[i] + P.ess_dof[i];
    P.prescribed_eq.push_back(eq);
    if (!P.lmm) P.ignore[eq] = 1;
  }
  P.ready = true;
  return nullptr;
}

// geometry at one IP: N, G (nnode×ndim), detJ, radius
struct Geo { double N[27]; double G[27 * 3]; double detJ; double r; };

const char* calc_geo(const Problem& P, int e, const IP& ip, Geo& g) {
  int kind = P.kind[e];
  const Basis& b = basis(kind);
  int nn = b.nnode, nd = P.ndim;
  double dN[27 * 3];
  eval_basis(b, ip.x, g.N, dN);
  const int* pts = &P.cpts[P.cptr[e]];
  double J[9] = {0};  // J[i][j] = dx_i/dxi_j
  for (int m = 0; m < nn; m++) for (int i = 0; i < nd; i++) for (int j = 0; j < nd; j++)
    J[i * 3 + j] += P.X[(size_t)pts[m] * nd + i] * dN[m * nd + j];
  double Ji[9];
  if (nd == 2) {
    g.detJ = J[0] * J[4] - J[1] * J[3];
    Ji[0] = J[4] / g.detJ; Ji[1] = -J[1] / g.detJ; Ji[3] = -J[3] / g.detJ; Ji[4] = J[0] / g.detJ;
  } else {
    double a = J[0], bb = J[1], c = J[2], d = J[3], ee = J[4], f = J[5], gg = J[6], h = J[7], k = J[8];
    g.detJ = a * (ee * k - f * h) - bb * (d * k - f * gg) + c * (d * h - ee * gg);
    double id = 1.0 / g.detJ;
    Ji[0] = (ee * k - f * h) * id; Ji[1] = (c * h - bb * k) * id; Ji[2] = (bb * f - c * ee) * id;
    Ji[3] = (f * gg - d * k) * id; Ji[4] = (a * k - c * gg) * id; Ji[5] = (c * d - a * f) * id;
    Ji[6] = (d * h - ee * gg) * id; Ji[7] = (bb * gg - a * h) * id; Ji[8] = (a * ee - bb * d) * id;
  }
  if (g.detJ <= 0.0) return "negative Jacobian determinant";
  // G = dN/dx = dN/dxi * J^{-1}  (G[m][j] = sum_k dN[m][k] Ji[k][j])
  for (int m = 0; m < nn; m++) for (int j = 0; j < nd; j++) {
    double s = 0;
    for (int k = 0; k < nd; k++) s += dN[m * nd + k] * Ji[k * 3 + j];
    g.G[m * nd + j] = s;
  }
  g.r = 0.0;
  if (P.ideal.axisym) for (int m = 0; m < nn; m++) g.r += g.N[m] * P.X[(size_t)pts[m] * nd];
  return nullptr;
}

// Mandel B matrix (ns × nn*ndim)
void calc_B(const Problem& P, int nn, const Geo& g, double* B) {
  int nd = P.ndim, ns = P.ns, ne = nn * nd;
  std::fill(B, B + ns * ne, 0.0);
  for (int m = 0; m < nn; m++) {
    const double* G = &g.G[m * nd];
    if (nd == 2) {
      B[0 * ne + 2 * m] = G[0];
      B[1 * ne + 2 * m + 1] = G[1];
      if (P.ideal.axisym) B[2 * ne + 2 * m] = g.N[m] / g.r;
      B[3 * ne + 2 * m] = G[1] / SQ2; B[3 * ne + 2 * m + 1] = G[0] / SQ2;
    } else {
      B[0 * ne + 3 * m] = G[0]; B[1 * ne + 3 * m + 1] = G[1]; B[2 * ne + 3 * m + 2] = G[2];
      B[3 * ne + 3 * m] = G[1] / SQ2; B[3 * ne + 3 * m + 1] = G[0] / SQ2;
      B[4 * ne + 3 * m + 1] = G[2] / SQ2; B[4 * ne + 3 * m + 2] = G[1] / SQ2;
      B[5 * ne + 3 * m] = G[2] / SQ2; B[5 * ne + 3 * m + 2] = G[0] / SQ2;
    }
  }
}

double coef(const Problem& P, const IP& ip, const Geo& g) {
  double c = g.detJ * ip.w * P.ideal.thickness;
  if (P.ideal.axisym) c *= g.r;
  return c;
}

// calculate_strain.rs:18-60 (same per-node accumulation as the reference)
void calc_strain(const Problem& P, int e, const Geo& g, const double* uu, double* eps) {
  int nd = P.ndim, nn = basis(P.kind[e]).nnode;
  const int* pts = &P.cpts[P.cptr[e]];
  for (int i = 0; i < 6; i++) eps[i] = 0;
  // tensor-form accumulation, then Mandel
  double exx = 0, eyy = 0, ezz = 0, exy = 0, eyz = 0, exz = 0;
  for (int m = 0; m < nn; m++) {
    const double* G = &g.G[m * nd];
    size_t b = (size_t)pts[m] * nd;
    if (nd == 2) {
      double ux = uu[b], uy = uu[b + 1];
      exx += ux * G[0]; eyy += uy * G[1]; exy += (ux * G[1] + uy * G[0]) / 2.0;
    } else {
      double ux = uu[b], uy = uu[b + 1], uz = uu[b + 2];
      exx += ux * G[0]; eyy += uy * G[1]; ezz += uz * G[2];
      exy += (ux * G[1] + uy * G[0]) / 2.0; eyz += (uy * G[2] + uz * G[1]) / 2.0; exz += (ux * G[2] + uz * G[0]) / 2.0;
    }
  }
  if (nd == 2 && P.ideal.axisym) for (int m = 0; m < nn; m++) ezz += uu[(size_t)pts[m] * nd] * g.N[m] / g.r;
  eps[0] = exx; eps[1] = eyy; eps[2] = ezz; eps[3] = exy * SQ2;
  if (nd == 3) { eps[4] = eyz * SQ2; eps[5] = exz * SQ2; }
}

// element_solid.rs:190-224 for every cell
const char* update_secondary(Problem& P, const double* ddu) {
  const char* err = nullptr;
#pragma omp parallel for schedule(static)
  for (int e = 0; e < P.ncell; e++) {
    const Param& pa = P.params.at(P.attr[e]);
    for (int ip = 0; ip < (int)P.rules[e].size(); ip++) {
      Geo g; const char* ge = calc_geo(P, e, P.rules[e][ip], g);
      double eps[6];
      if (ge) { err = ge; continue; }
      calc_strain(P, e, g, ddu, eps);
      IPState& s = P.st[P.ip_ptr[e] + ip];
      const char* ue = update_stress(pa, P.ideal, P.ns, s, eps);
      if (ue) err = ue;
    }
  }
  return err;
}

// local f_int (element_solid.rs:124-138 -> vec_04_tb)
void calc_f_int(const Problem& P, int e, double* f) {
  int nd = P.ndim, nn = basis(P.kind[e]).nnode;
  for (int i = 0; i < nn * nd; i++) f[i] = 0;
  for (int ip = 0; ip < (int)P.rules[e].size(); ip++) {
    Geo g; calc_geo(P, e, P.rules[e][ip], g);
    double c = coef(P, P.rules[e][ip], g);
    const double* s = P.st[P.ip_ptr[e] + ip].sig;
    for (int m = 0; m < nn; m++) {
      const double* G = &g.G[m * nd];
      if (nd == 2) {
        double sxy = s[3] / SQ2;
        double fx = s[0] * G[0] + sxy * G[1];
        if (P.ideal.axisym) fx += s[2] * g.N[m] / g.r;
        f[2 * m] += c * fx;
        f[2 * m + 1] += c * (sxy * G[0] + s[1] * G[1]);
      } else {
        double sxy = s[3] / SQ2, syz = s[4] / SQ2, sxz = s[5] / SQ2;
        f[3 * m] += c * (s[0] * G[0] + sxy * G[1] + sxz * G[2]);
        f[3 * m + 1] += c * (sxy * G[0] + s[1] * G[1] + syz * G[2]);
        f[3 * m + 2] += c * (sxz * G[0] + syz * G[1] + s[2] * G[2]);
      }
    }
  }
}

// local K (element_solid.rs:168-185 -> mat_10_bdb): dense B^T (D B) per IP
const char* calc_K(const Problem& P, int e, double* K) {
  int nd = P.ndim, nn = basis(P.kind[e]).nnode, ne = nn * nd, ns = P.ns;
  const Param& pa = P.params.at(P.attr[e]);
  std::fill(K, K + ne * ne, 0.0);
  double B[6 * 60], DB[6 * 60];
  for (int ip = 0; ip < (int)P.rules[e].size(); ip++) {
    Geo g; calc_geo(P, e, P.rules[e][ip], g);
    double c = coef(P, P.rules[e][ip], g);
    double D[36];
    const char* se = stiffness(pa, P.ideal, ns, P.st[P.ip_ptr[e] + ip], D);
    if (se) return se;
    calc_B(P, nn, g, B);
    for (int i = 0; i < ns; i++) for (int j = 0; j < ne; j++) {
      double s = 0; for (int k = 0; k < ns; k++) s += D[i * ns + k] * B[k * ne + j];
      DB[i * ne + j] = s;
    }
    for (int i = 0; i < ne; i++) for (int j = 0; j < ne; j++) {
      double s = 0; for (int k = 0; k < ns; k++) s += B[k * ne + i] * DB[k * ne + j];
      K[i * ne + j] += c * s;
    }
  }
  return nullptr;
}

// boundary facet load (BcDistributed::calc_f_ext, bc_distributed.rs:156-236): Qn: f = int N qn n dGamma with
// the outward unit normal n (bc_distributed.rs:182-187); Qx/Qy/Qz: f = int N q e_k dGamma (bc_distributed.rs:195-231)
void add_facet_load(const Problem& P, const Facet& F, double* ff) {
  const double value = load_value(P, F.values.data(), (int)F.values.size(), P.t);
  const Basis& b = basis(F.kind);
  int nn = b.nnode, nd = P.ndim;
  std::vector<IP> rule;
  if (b.family == 4) rule = gauss_legendre_1d(std::max(3, nn));
  else if (b.family == 0) rule = tri_rule(nn <= 3 ? 3 : 6);
  else rule = gauss_rule(F.kind, 9);
  // cell centroid (outward orientation of Qn)
  double cc[3] = {0, 0, 0};
  if (F.cell >= 0) {
    int cn = P.cptr[F.cell + 1] - P.cptr[F.cell];
    for (int k = 0; k < cn; k++) for (int i = 0; i < nd; i++) cc[i] += P.X[(size_t)P.cpts[P.cptr[F.cell] + k] * nd + i] / cn;
  }
  double fc[3] = {0, 0, 0};
  for (int k = 0; k < nn; k++) for (int i = 0; i < nd; i++) fc[i] += P.X[(size_t)F.pts[k] * nd + i] / nn;
  const bool line = b.family == 4;
  for (auto& ip : rule) {
    double N[27], dN[27 * 3];
    eval_basis(b, ip.x, N, dN);
    // area vector: a normal whose length is the measure of the facet per unit reference measure; for a
    // line in 3D, its tangent (only its length is used)
    double n[3] = {0, 0, 0};
    if (nd == 2) {
      double dx = 0, dy = 0;
      for (int m = 0; m < nn; m++) { dx += P.X[(size_t)F.pts[m] * 2] * dN[m]; dy += P.X[(size_t)F.pts[m] * 2 + 1] * dN[m]; }
      n[0] = dy; n[1] = -dx;
    } else if (line) {
      for (int m = 0; m < nn; m++) for (int i = 0; i < 3; i++) n[i] += P.X[(size_t)F.pts[m] * 3 + i] * dN[m];
    } else {
      double a[3] = {0, 0, 0}, c[3] = {0, 0, 0};
      for (int m = 0; m < nn; m++) for (int i = 0; i < 3; i++) {
        a[i] += P.X[(size_t)F.pts[m] * 3 + i] * dN[m * 2];
        c[i] += P.X[(size_t)F.pts[m] * 3 + i] * dN[m * 2 + 1];
      }
      n[0] = a[1] * c[2] - a[2] * c[1]; n[1] = a[2] * c[0] - a[0] * c[2]; n[2] = a[0] * c[1] - a[1] * c[0];
    }
    double w = ip.w * (nd == 2 ? P.ideal.thickness : 1.0);
    if (nd == 2 && P.ideal.axisym) {
      double r = 0; for (int m = 0; m < nn; m++) r += N[m] * P.X[(size_t)F.pts[m] * 2];
      w *= r;
    }
    double v[3] = {0, 0, 0};
    if (F.nbc == 0) {
      double dot = 0;
      for (int i = 0; i < nd; i++) dot += n[i] * (fc[i] - cc[i]);
      double sgn = dot < 0 ? -1.0 : 1.0;
      for (int i = 0; i < nd; i++) v[i] = value * sgn * n[i];
    } else {
      const double len = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
      v[F.nbc - 1] = value * len;
    }
    for (int m = 0; m < nn; m++) for (int i = 0; i < nd; i++) ff[(size_t)F.pts[m] * nd + i] += N[m] * v[i] * w;
  }
}

// ---------------------------------------------------------------------------------------------
// banded LU with partial pivoting (for the literal LMM/RSM global system)
// ---------------------------------------------------------------------------------------------
struct Band {
  int n = 0, kl = 0, ku = 0, w = 0;  // storage row i: columns [i-kl, i-kl+w)
  std::vector<double> a;
  std::vector<int> piv;
  double& at(int i, int j) { return a[(size_t)i * w + (j - i + kl)]; }
};

// factorization in place: U in the upper band, the multipliers of column k below its diagonal, row
// interchanges in piv (applied to the right-hand side in elimination order, as LAPACK's gbtrf/gbtrs)
bool band_lu_factor(Band& B) {
  int n = B.n, kl = B.kl;
  int ucols = B.ku + kl;  // upper extent after pivoting
  B.piv.assign(n, 0);
  for (int k = 0; k < n; k++) {
    int last = std::min(n - 1, k + kl);
    int p = k; double best = std::fabs(B.at(k, k));
    for (int i = k + 1; i <= last; i++) if (std::fabs(B.at(i, k)) > best) { best = std::fabs(B.at(i, k)); p = i; }
    if (best == 0.0) return false;
    B.piv[k] = p;
    int jend = std::min(n - 1, k + ucols);
    if (p != k) for (int j = k; j <= jend; j++) std::swap(B.at(k, j), B.at(p, j));
    double d = B.at(k, k);
    for (int i = k + 1; i <= last; i++) {
      double f = B.at(i, k) / d;
      B.at(i, k) = f;
      if (f == 0.0) continue;
      for (int j = k + 1; j <= jend; j++) B.at(i, j) -= f * B.at(k, j);
    }
  }
  return true;
}

void band_lu_apply(Band& B, std::vector<double>& x) {
  int n = B.n, kl = B.kl, ucols = B.ku + kl;
  for (int k = 0; k < n; k++) {
    if (B.piv[k] != k) std::swap(x[k], x[B.piv[k]]);
    int last = std::min(n - 1, k + kl);
    for (int i = k + 1; i <= last; i++) x[i] -= B.at(i, k) * x[k];
  }
  for (int k = n - 1; k >= 0; k--) {
    double s = x[k];
    int jend = std::min(n - 1, k + ucols);
    for (int j = k + 1; j <= jend; j++) s -= B.at(k, j) * x[j];
    x[k] = s / B.at(k, k);
  }
}

// direct solve with two steps of iterative refinement against the unfactored matrix A (as UMFPACK's
// default, UMFPACK_IRSTEP = 2, used by pmsim's default Genie; src/fem/linear_system.rs:185-201)
void band_lu_refined(const Band& A, Band& B, std::vector<double>& x) {  // B: factors of A
  const std::vector<double> b = x;
  band_lu_apply(B, x);
  std::vector<double> r(B.n);
  for (int step = 0; step < 2; step++) {
    for (int i = 0; i < A.n; i++) {
      double s = b[i];
      const int j0 = std::max(0, i - A.kl), j1 = std::min(A.n - 1, i + A.kl);
      for (int j = j0; j <= j1; j++) s -= A.a[(size_t)i * A.w + (j - i + A.kl)] * x[j];
      r[i] = s;
    }
    band_lu_apply(B, r);
    for (int i = 0; i < A.n; i++) x[i] += r[i];
  }
}

bool band_lu_solve(Band& B, std::vector<double>& x) {
  const Band A = B;
  if (!band_lu_factor(B)) return false;
  band_lu_refined(A, B, x);
  return true;
}

// reverse Cuthill-McKee over points
std::vector<int> rcm_points(const Problem& P) {
  std::vector<std::vector<int>> adj(P.npoint);
  for (int e = 0; e < P.ncell; e++) {
    for (int a = P.cptr[e]; a < P.cptr[e + 1]; a++) for (int b = P.cptr[e]; b < P.cptr[e + 1]; b++)
      if (a != b) adj[P.cpts[a]].push_back(P.cpts[b]);
  }
  for (auto& v : adj) { std::sort(v.begin(), v.end()); v.erase(std::unique(v.begin(), v.end()), v.end()); }
  std::vector<int> order; order.reserve(P.npoint);
  std::vector<char> seen(P.npoint, 0);
  while ((int)order.size() < P.npoint) {
    int start = -1;
    for (int i = 0; i < P.npoint; i++) if (!seen[i] && (start < 0 || adj[i].size() < adj[start].size())) start = i;
    size_t head = order.size();
    order.push_back(start); seen[start] = 1;
    while (head < order.size()) {
      int v = order[head++];
      std::vector<int> nb;
      for (int w : adj[v]) if (!seen[w]) { nb.push_back(w); seen[w] = 1; }
      std::sort(nb.begin(), nb.end(), [&](int a, int b) { return adj[a].size() < adj[b].size() || (adj[a].size() == adj[b].size() && a < b); });
      for (int w : nb) order.push_back(w);
    }
  }
  std::reverse(order.begin(), order.end());
  return order;
}

// ---------------------------------------------------------------------------------------------
// Newton driver (solver_implicit.rs:56-272)
// ---------------------------------------------------------------------------------------------
struct Out { double* u; double* ipv; int max_out; int n_out; double* ell = nullptr; };

void write_output(Problem& P, Out& o) {
  if (o.n_out >= o.max_out) { o.n_out++; return; }
  if (o.ell) o.ell[o.n_out] = P.ell;
  if (o.u) std::memcpy(o.u + (size_t)o.n_out * P.neq_total, P.u.data(), sizeof(double) * P.neq_total);
  if (o.ipv) {
    int nv = P.ns + 4;
    double* dst = o.ipv + (size_t)o.n_out * P.nip_total * nv;
    for (int i = 0; i < P.nip_total; i++) {
      const IPState& s = P.st[i];
      for (int k = 0; k < P.ns; k++) dst[i * nv + k] = s.sig[k];
      dst[i * nv + P.ns] = s.z; dst[i * nv + P.ns + 1] = s.lam;
      dst[i * nv + P.ns + 2] = s.elastic ? 1.0 : 0.0; dst[i * nv + P.ns + 3] = s.epbar;
    }
  }
  o.n_out++;
}

struct Sys {
  std::vector<int> perm, iperm;   // eq -> position
  int kl = 0;
};

Sys build_ordering(Problem& P) {
  Sys S;
  int nt = P.neq_total;
  auto ord = rcm_points(P);
  std::vector<std::vector<int>> lam_of_eq(P.ndof);
  if (P.lmm) for (size_t p = 0; p < P.prescribed_eq.size(); p++) lam_of_eq[P.prescribed_eq[p]].push_back(P.ndof + (int)p);
  S.perm.clear();
  for (int pt : ord) for (int d = 0; d < P.ndim; d++) {
    int eq = P.ndim * pt + d;
    S.perm.push_back(eq);
    for (int l : lam_of_eq[eq]) S.perm.push_back(l);
  }
  S.iperm.assign(nt, 0);
  for (int i = 0; i < nt; i++) S.iperm[S.perm[i]] = i;
  // bandwidth from element connectivity (+ LMM couplings)
  int bw = 0;
  for (int e = 0; e < P.ncell; e++) {
    int lo = 1 << 30, hi = -1;
    for (int a = P.cptr[e]; a < P.cptr[e + 1]; a++) for (int d = 0; d < P.ndim; d++) {
      int pos = S.iperm[P.ndim * P.cpts[a] + d]; lo = std::min(lo, pos); hi = std::max(hi, pos);
    }
    bw = std::max(bw, hi - lo);
  }
  if (P.lmm) for (size_t p = 0; p < P.prescribed_eq.size(); p++)
    bw = std::max(bw, std::abs(S.iperm[P.prescribed_eq[p]] - S.iperm[P.ndof + (int)p]));
  S.kl = bw;
  return S;
}

const char* solve(Problem& P, Out& o) {
  if (const char* e = setup(P)) return e;
  // has_non_zero check (bc_prescribed.rs:74-81; the constant is the table value scale)
  if (!P.lmm) for (double v : P.ess_table) if (v != 0.0) return "the Lagrange multiplier method is required for non-zero prescribed values";
  int nt = P.neq_total, ne_max = 60;
  (void)ne_max;
  P.iters.clear(); P.norm_rr_hist.clear();
  // config.validate (config.rs:255-265), reported as solver_data.rs:45-47 does
  if (P.arc_psi < 0.0 || P.arc_psi > 1.0 || std::fabs(P.ell0) < 1e-12)
    return "cannot allocate simulation because config.validate() failed";
  P.ell = 0.0;
  // time_control.initialize
  P.t = P.t_ini;
  if (P.dt < 1e-10) return "Δt is smaller than the allowed minimum";
  // initialize_internal_values (von_mises.rs:107-116)
  for (int e = 0; e < P.ncell; e++) {
    const Param& pa = P.params.at(P.attr[e]);
    for (int ip = P.ip_ptr[e]; ip < P.ip_ptr[e + 1]; ip++) {
      IPState& s = P.st[ip];
      if (pa.model == 1) { s.z = pa.z0; if (sigma_d(s.sig, P.ns) - s.z > 0) return "stress is outside the yield surface"; }
    }
  }
  write_output(P, o);
  double t_out = P.t + P.dt_out;
  Sys S = build_ordering(P);
  std::vector<double> ff_int(nt), ff_ext(nt), rr(nt), mdu(nt), mdu0(nt);
  int nd = P.ndim;
  int step_n_failure = 0;
  bool converged = false;
  // arc-length control (arc_length_control.rs:7-48): Δs, its bounds, ℓ and u of the two previous
  // converged steps, the constraint g and its derivatives, the iterative δℓ and the total Δℓ
  double dds = 0, dds_old = 0, dds_min = 0, dds_max = 0, ell_anc = 0, ell_old = 0, arc_g = 0, dg_dl = 0, ddl = 0, arc_dl = 0;
  std::vector<double> u_anc(nt, 0.0), u_old(nt, 0.0), ddu1(nt), ddu2(nt);
  bool converged_old = false;
  auto dot = [&](const std::vector<double>& a, const std::vector<double>& b) {
    double r = 0; for (int i = 0; i < nt; i++) r += a[i] * b[i]; return r;
  };
  for (int timestep = 0; timestep < P.n_max_time_steps; timestep++) {
    // time_control.update
    if (P.t + P.dt > P.t_fin) { write_output(P, o); break; }
    P.t += P.dt;
    if (P.arc) {
      // trial_increments (arc_length_control.rs:56-83): extrapolate u and ℓ from the two previous
      // converged steps scaled by Δs/Δs_old; Δu = u - u_old, Δℓ = ℓ - ℓ_old
      if (timestep == 0) P.ell = P.ell0;
      if (timestep > 0) {
        if (std::fabs(dds_old) < 1e-12) return "Δs_old is too small";
        const double alpha = dds / dds_old;
        for (int i = 0; i < nt; i++) P.u[i] = (1.0 + alpha) * u_old[i] - alpha * u_anc[i];
        P.ell = (1.0 + alpha) * ell_old - alpha * ell_anc;
      }
      for (int i = 0; i < nt; i++) P.ddu[i] = P.u[i] - u_old[i];
      ddl = P.ell - ell_old;
      converged_old = converged;
    } else {
      std::fill(P.ddu.begin(), P.ddu.end(), 0.0);
      P.ell = 1.0;
    }
    converged = false;
    if (!P.linear_problem) for (auto& s : P.st) s.lam = 0.0;  // reset_algorithmic_variables
    int its = 0;
    for (int iteration = 0; iteration < P.n_max_iterations; iteration++) {
      its = iteration + 1;
      // ---- iterate() ----
      std::fill(ff_int.begin(), ff_int.end(), 0.0);
      std::fill(ff_ext.begin(), ff_ext.end(), 0.0);
      std::vector<double> fe(60);
      for (int e = 0; e < P.ncell; e++) {
        calc_f_int(P, e, fe.data());
        int nn = P.cptr[e + 1] - P.cptr[e];
        for (int m = 0; m < nn; m++) for (int d = 0; d < nd; d++) {
          int g = nd * P.cpts[P.cptr[e] + m] + d;
          if (!P.ignore[g]) ff_int[g] += fe[nd * m + d];
        }
      }
      {
        std::vector<double> fx(P.ndof, 0.0);
        for (auto& F : P.facets) add_facet_load(P, F, fx.data());
        for (int g = 0; g < P.ndof; g++) if (!P.ignore[g]) ff_ext[g] += fx[g];
        for (size_t k = 0; k < P.pl_point.size(); k++)
          ff_ext[nd * P.pl_point[k] + P.pl_dof[k]] += load_value(P, &P.pl_table[k * P.pl_nstation], P.pl_nstation, P.t);
      }
      for (int i = 0; i < nt; i++) rr[i] = ff_int[i] - P.ell * ff_ext[i];
      if (P.lmm) for (size_t p = 0; p < P.prescribed_eq.size(); p++) {
        int i = P.prescribed_eq[p], j = P.ndof + (int)p;
        rr[i] += P.u[j];
        rr[j] = P.u[i] - table_value(P, (int)p, P.t);
      }
      // arc-length constraint g = Δu·Δu + ψ Δℓ² F·F - Δs² and its derivatives (arc_length_control.rs:86-105)
      arc_g = 0.0;
      if (P.arc) {
        if (timestep > 0) {
          const double ftf = dot(ff_ext, ff_ext);
          arc_g = dot(P.ddu, P.ddu) + P.arc_psi * ddl * ddl * ftf - dds * dds;
          dg_dl = 2.0 * P.arc_psi * ddl * ftf;
        } else {
          dg_dl = 1.0;
        }
      }
      // analyze_rr: max(|R|inf, |g|) (convergence_control.rs:52-67)
      double norm_rr = std::fabs(arc_g); bool bad = !std::isfinite(arc_g);
      for (double v : rr) { if (!std::isfinite(v)) bad = true; norm_rr = std::max(norm_rr, std::fabs(v)); }
      if (bad) return "Found NaN or Inf";
      P.norm_rr_hist.push_back(norm_rr);
      if (norm_rr < P.tol_rr_abs) { converged = true; break; }
      // assemble K in band storage
      Band B; B.n = nt; B.kl = S.kl; B.ku = S.kl; B.w = 2 * S.kl + S.kl + 1;
      B.a.assign((size_t)nt * B.w, 0.0);
      std::vector<double> Ke(60 * 60);
      for (int e = 0; e < P.ncell; e++) {
        const char* ke = calc_K(P, e, Ke.data());
        if (ke) return ke;
        int nn = P.cptr[e + 1] - P.cptr[e], nl = nn * nd;
        std::vector<int> l2g(nl);
        for (int m = 0; m < nn; m++) for (int d = 0; d < nd; d++) l2g[nd * m + d] = nd * P.cpts[P.cptr[e] + m] + d;
        for (int l = 0; l < nl; l++) {
          int g = l2g[l]; if (P.ignore[g]) continue;
          for (int ll = 0; ll < nl; ll++) {
            int gg = l2g[ll]; if (P.ignore[gg]) continue;
            B.at(S.iperm[g], S.iperm[gg]) += Ke[l * nl + ll];
          }
        }
      }
      if (P.lmm) {
        for (size_t p = 0; p < P.prescribed_eq.size(); p++) {
          int i = S.iperm[P.prescribed_eq[p]], j = S.iperm[P.ndof + (int)p];
          B.at(i, j) += 1.0; B.at(j, i) += 1.0;
        }
      } else {
        for (int eq : P.prescribed_eq) B.at(S.iperm[eq], S.iperm[eq]) += 1.0;
      }
      std::vector<double> x(nt);
      if (P.arc) {
        // arc-length solve (arc_length_control.rs:108-126): Δu1 = K⁻¹ F_ext, Δu2 = K⁻¹ R,
        // δℓ = (dg/du·Δu2 - g) / (dg/dℓ + dg/du·Δu1), -δu = Δu2 - δℓ Δu1, with dg/du = 2Δu (0 at step 0)
        const Band A = B;
        if (!band_lu_factor(B)) return "linear solver failed (singular matrix)";
        for (int i = 0; i < nt; i++) x[S.iperm[i]] = ff_ext[i];
        band_lu_refined(A, B, x);
        for (int i = 0; i < nt; i++) ddu1[i] = x[S.iperm[i]];
        for (int i = 0; i < nt; i++) x[S.iperm[i]] = rr[i];
        band_lu_refined(A, B, x);
        for (int i = 0; i < nt; i++) ddu2[i] = x[S.iperm[i]];
        double c1 = 0, c2 = 0;
        if (timestep > 0) for (int i = 0; i < nt; i++) { c1 += 2.0 * P.ddu[i] * ddu1[i]; c2 += 2.0 * P.ddu[i] * ddu2[i]; }
        arc_dl = (c2 - arc_g) / (dg_dl + c1);
        for (int i = 0; i < nt; i++) mdu[i] = ddu2[i] - arc_dl * ddu1[i];
      } else {
        for (int i = 0; i < nt; i++) x[S.iperm[i]] = rr[i];
        if (!band_lu_solve(B, x)) return "linear solver failed (singular matrix)";
        for (int i = 0; i < nt; i++) mdu[i] = x[S.iperm[i]];
      }
      // analyze_mdu
      double norm_mdu = 0; bad = false;
      for (double v : mdu) { if (!std::isfinite(v)) bad = true; norm_mdu = std::max(norm_mdu, std::fabs(v)); }
      if (bad) return "Found NaN or Inf in mdu";
      bool conv_mdu = false;
      if (iteration == 0) mdu0 = mdu;
      else {
        double rel = 0;
        for (int i = 0; i < nt; i++) rel = std::max(rel, std::fabs(mdu[i]) / (1.0 + std::fabs(mdu0[i])));
        conv_mdu = rel < P.tol_mdu_rel;
      }
      if (conv_mdu) { converged = true; break; }
      // update primary variables (unknown equations)
      for (int i = 0; i < nt; i++) {
        if (i < P.ndof && P.ignore[i]) continue;
        P.u[i] -= mdu[i]; P.ddu[i] -= mdu[i];
      }
      if (P.arc) { P.ell += arc_dl; ddl += arc_dl; }  // update_load_factor: ℓ += δℓ, Δℓ += δℓ
      if (!P.linear_problem) {
        if (iteration == 0) P.backup = P.st; else P.st = P.backup;
      }
      // update secondary values; epbar tracks the step-start value + current lambda
      if (const char* ue = update_secondary(P, P.ddu.data())) return ue;
      for (int i = 0; i < P.nip_total; i++) P.st[i].epbar = P.backup[i].epbar + P.st[i].lam;
      if (P.linear_problem) { converged = true; break; }
      if (!P.arc && iteration == P.n_max_iterations - 1) return "Newton-Raphson did not converge";
    }
    if (!converged && !P.arc) return "Newton-Raphson did not converge";
    P.iters.push_back(converged ? its : -its);
    if (P.arc) {
      // step_adaptation (arc_length_control.rs:136-168)
      if (converged) {
        if (timestep == 0) {
          dds = std::sqrt(dot(P.ddu, P.ddu) + P.arc_psi * ddl * ddl * dot(ff_ext, ff_ext));
          dds_max = dds;
          dds_min = dds / 1024.0;
        }
        dds_old = dds;
        if (converged_old) dds = std::min(std::max(2.0 * dds, dds_min), dds_max);
        u_anc = u_old; u_old = P.u;
        ell_anc = ell_old; ell_old = P.ell;
      } else {
        dds = converged_old ? std::max(dds / 2.0, dds_min) : std::max(dds / 4.0, dds_min);
      }
    }
    bool last = timestep == P.n_max_time_steps - 1;
    if ((converged && P.t >= t_out) || last) { write_output(P, o); t_out += P.dt_out; }
    step_n_failure++;
    if (step_n_failure > P.allowed_step_n_failure) return "too many step failures";
    if (P.t >= P.t_fin) break;
  }
  return nullptr;
}

}  // namespace

// =============================================================================================
// C API (ctypes)
// =============================================================================================
extern "C" {

void* orc_new(int ndim, int npoint, const double* coords, int ncell, const int* kinds, const int* attrs,
              const int* cell_ptr, const int* cell_pts) {
  auto* P = new Problem();
  P->ndim = ndim; P->npoint = npoint; P->ncell = ncell;
  P->X.assign(coords, coords + (size_t)npoint * ndim);
  P->kind.assign(kinds, kinds + ncell);
  P->attr.assign(attrs, attrs + ncell);
  P->cptr.assign(cell_ptr, cell_ptr + ncell + 1);
  P->cpts.assign(cell_pts, cell_pts + cell_ptr[ncell]);
  return P;
}

void orc_free(void* h) { delete (Problem*)h; }

const char* orc_last_error(void* h) { return ((Problem*)h)->err.c_str(); }

int orc_set_param(void* h, int attr, int model, double E, double nu, double H, double z0, int ngauss) {
  Param p; p.model = model; p.E = E; p.nu = nu; p.H = H; p.z0 = z0; p.ngauss = ngauss;
  ((Problem*)h)->params[attr] = p;
  return 0;
}

int orc_set_ideal(void* h, int axisym, int plane_stress, double thickness) {
  auto* P = (Problem*)h;
  P->ideal.ndim = P->ndim; P->ideal.axisym = axisym; P->ideal.plane_stress = plane_stress; P->ideal.thickness = thickness;
  return 0;
}

int orc_set_config(void* h, int lmm, int linear_problem, double t_ini, double t_fin, double dt, int n_max_iterations,
                   double tol_rr_abs, double tol_mdu_rel) {
  auto* P = (Problem*)h;
  P->lmm = lmm; P->linear_problem = linear_problem; P->t_ini = t_ini; P->t_fin = t_fin; P->dt = dt; P->dt_out = dt;
  P->n_max_iterations = n_max_iterations; P->tol_rr_abs = tol_rr_abs; P->tol_mdu_rel = tol_mdu_rel;
  return 0;
}

int orc_set_essential(void* h, int n, const int* pts, const int* dofs, const double* table, int nstation) {
  auto* P = (Problem*)h;
  P->ess_point.assign(pts, pts + n); P->ess_dof.assign(dofs, dofs + n);
  P->ess_table.assign(table, table + (size_t)n * nstation); P->nstation = nstation;
  return 0;
}

int orc_set_point_loads(void* h, int n, const int* pts, const int* dofs, const double* table, int nstation) {
  auto* P = (Problem*)h;
  P->pl_point.assign(pts, pts + n); P->pl_dof.assign(dofs, dofs + n);
  P->pl_table.assign(table, table + (size_t)n * nstation); P->pl_nstation = nstation;
  return 0;
}

int orc_add_facet_qn(void* h, int facet_kind, int npts, const int* pts, int cell, const double* values, int nstation) {
  auto* P = (Problem*)h;
  P->facets.push_back({facet_kind, std::vector<int>(pts, pts + npts), cell, std::vector<double>(values, values + nstation)});
  return 0;
}

int orc_add_facet_load(void* h, int nbc, int facet_kind, int npts, const int* pts, int cell, const double* values, int nstation) {
  auto* P = (Problem*)h;
  Facet F{facet_kind, std::vector<int>(pts, pts + npts), cell, std::vector<double>(values, values + nstation)};
  F.nbc = nbc;
  P->facets.push_back(F);
  return 0;
}

// F_ext at time t (solver_data.rs:85-108): distributed loads (ignored equations skipped), then point loads
int orc_f_ext(void* h, double t, double* out) {
  auto* P = (Problem*)h;
  const double t_keep = P->t;
  P->t = t;
  std::vector<double> fx(P->ndof, 0.0);
  for (auto& F : P->facets) add_facet_load(*P, F, fx.data());
  for (int g = 0; g < P->ndof; g++) out[g] = P->ignore[g] ? 0.0 : fx[g];
  for (size_t k = 0; k < P->pl_point.size(); k++)
    out[P->ndim * P->pl_point[k] + P->pl_dof[k]] += load_value(*P, &P->pl_table[k * P->pl_nstation], P->pl_nstation, P->t);
  P->t = t_keep;
  return 0;
}

int orc_setup(void* h) {
  auto* P = (Problem*)h;
  if (const char* e = setup(*P)) { P->err = e; return 1; }
  return 0;
}

int orc_sizes(void* h, int* out) {  // [ndof, neq_total, nip_total, nsig]
  auto* P = (Problem*)h;
  if (const char* e = setup(*P)) { P->err = e; return 1; }
  out[0] = P->ndof; out[1] = P->neq_total; out[2] = P->nip_total; out[3] = P->ns;
  return 0;
}

// full Newton solve; u_out: (max_out × neq_total), ipv_out: (max_out × nip × (ns+4))
int orc_set_arc_length(void* h, int enable, double psi, double ell0) {
  auto* P = (Problem*)h;
  P->arc = enable != 0; P->arc_psi = psi; P->ell0 = ell0;
  return 0;
}

// ell_out (nullable): the loading factor of each output
int orc_solve(void* h, double* u_out, double* ipv_out, int max_out, int* n_out, int* iters_out, int max_steps, double* ell_out) {
  auto* P = (Problem*)h;
  Out o{u_out, ipv_out, max_out, 0, ell_out};
  P->ready = false;
  const char* e = solve(*P, o);
  *n_out = o.n_out;
  for (int i = 0; i < (int)P->iters.size() && i < max_steps; i++) iters_out[i] = P->iters[i];
  if (e) { P->err = e; return 1; }
  return 0;
}

// ---- element-level access (for element-pass parity) ----
int orc_init_state(void* h) {
  auto* P = (Problem*)h;
  if (const char* e = setup(*P)) { P->err = e; return 1; }
  for (int e = 0; e < P->ncell; e++) {
    const Param& pa = P->params.at(P->attr[e]);
    for (int ip = P->ip_ptr[e]; ip < P->ip_ptr[e + 1]; ip++) { P->st[ip] = IPState(); if (pa.model == 1) P->st[ip].z = pa.z0; }
  }
  return 0;
}

int orc_set_ip_state(void* h, const double* v) {  // nip × (ns+4)
  auto* P = (Problem*)h;
  int nv = P->ns + 4;
  for (int i = 0; i < P->nip_total; i++) {
    IPState& s = P->st[i];
    for (int k = 0; k < P->ns; k++) s.sig[k] = v[i * nv + k];
    s.z = v[i * nv + P->ns]; s.lam = v[i * nv + P->ns + 1]; s.elastic = v[i * nv + P->ns + 2] != 0.0; s.epbar = v[i * nv + P->ns + 3];
  }
  return 0;
}

int orc_get_ip_state(void* h, double* v) {
  auto* P = (Problem*)h;
  int nv = P->ns + 4;
  for (int i = 0; i < P->nip_total; i++) {
    const IPState& s = P->st[i];
    for (int k = 0; k < P->ns; k++) v[i * nv + k] = s.sig[k];
    v[i * nv + P->ns] = s.z; v[i * nv + P->ns + 1] = s.lam; v[i * nv + P->ns + 2] = s.elastic ? 1.0 : 0.0; v[i * nv + P->ns + 3] = s.epbar;
  }
  return 0;
}

// update from the current state with increment ddu (ndof); epbar += lambda
int orc_update_secondary(void* h, const double* ddu) {
  auto* P = (Problem*)h;
  std::vector<double> ep(P->nip_total);
  for (int i = 0; i < P->nip_total; i++) ep[i] = P->st[i].epbar;
  if (const char* e = update_secondary(*P, ddu)) { P->err = e; return 1; }
  for (int i = 0; i < P->nip_total; i++) P->st[i].epbar = ep[i] + P->st[i].lam;
  return 0;
}

// F_int over all DOFs (no ignore)
int orc_assemble_f_int(void* h, double* ff) {
  auto* P = (Problem*)h;
  int nd = P->ndim;
  std::fill(ff, ff + P->ndof, 0.0);
  std::vector<double> fe(60);
  for (int e = 0; e < P->ncell; e++) {
    calc_f_int(*P, e, fe.data());
    int nn = P->cptr[e + 1] - P->cptr[e];
    for (int m = 0; m < nn; m++) for (int d = 0; d < nd; d++) ff[nd * P->cpts[P->cptr[e] + m] + d] += fe[nd * m + d];
  }
  return 0;
}

// local stiffness of one cell (ne×ne)
int orc_element_K(void* h, int cell, double* K) {
  auto* P = (Problem*)h;
  if (const char* e = calc_K(*P, cell, K)) { P->err = e; return 1; }
  return 0;
}

int orc_element_f(void* h, int cell, double* f) {
  calc_f_int(*(Problem*)h, cell, f);
  return 0;
}

// global K in CSR (duplicates summed in element-ascending insertion order, as COO put + convert)
// first call with vals == nullptr returns nnz in *nnz and fills rowptr/cols
int orc_assemble_kk_csr(void* h, int64_t* rowptr, int* cols, double* vals, int64_t* nnz) {
  auto* P = (Problem*)h;
  int nd = P->ndim, n = P->ndof;
  std::vector<std::vector<int>> rows(n);
  for (int e = 0; e < P->ncell; e++) {
    int nn = P->cptr[e + 1] - P->cptr[e];
    for (int a = 0; a < nn; a++) for (int b = 0; b < nn; b++) for (int i = 0; i < nd; i++) for (int j = 0; j < nd; j++)
      rows[nd * P->cpts[P->cptr[e] + a] + i].push_back(nd * P->cpts[P->cptr[e] + b] + j);
  }
  int64_t k = 0;
  rowptr[0] = 0;
  for (int r = 0; r < n; r++) {
    auto& v = rows[r]; std::sort(v.begin(), v.end()); v.erase(std::unique(v.begin(), v.end()), v.end());
    for (int c : v) cols[k++] = c;
    rowptr[r + 1] = k;
  }
  *nnz = k;
  if (!vals) return 0;
  std::fill(vals, vals + k, 0.0);
  std::vector<double> Ke(60 * 60);
  for (int e = 0; e < P->ncell; e++) {
    if (const char* er = calc_K(*P, e, Ke.data())) { P->err = er; return 1; }
    int nn = P->cptr[e + 1] - P->cptr[e], nl = nn * nd;
    for (int l = 0; l < nl; l++) {
      int g = nd * P->cpts[P->cptr[e] + l / nd] + l % nd;
      for (int ll = 0; ll < nl; ll++) {
        int gg = nd * P->cpts[P->cptr[e] + ll / nd] + ll % nd;
        int64_t lo = rowptr[g], hi = rowptr[g + 1];
        int64_t pos = std::lower_bound(cols + lo, cols + hi, gg) - cols;
        vals[pos] += Ke[l * nl + ll];
      }
    }
  }
  return 0;
}

// Timed CPU element pass over all cells (update + f_int + K), used as the CPU baseline.
// Returns seconds; K values are accumulated into per-element buffers (not a global matrix).
double orc_time_element_pass(void* h, const double* ddu, int reps, int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
  auto* P = (Problem*)h;
  int nd = P->ndim;
  std::vector<IPState> saved = P->st;
  double sink = 0;
  auto t0 = std::chrono::steady_clock::now();
  for (int r = 0; r < reps; r++) {
    P->st = saved;
    update_secondary(*P, ddu);
    double local = 0;
#pragma omp parallel for schedule(static) reduction(+ : local)
    for (int e = 0; e < P->ncell; e++) {
      double fe[60]; double Ke[3600];
      calc_f_int(*P, e, fe);
      calc_K(*P, e, Ke);
      int nl = (P->cptr[e + 1] - P->cptr[e]) * nd;
      local += fe[0] + Ke[nl * nl - 1];
    }
    sink += local;
  }
  auto t1 = std::chrono::steady_clock::now();
  P->st = saved;
  if (sink == 12345.678) std::printf("%g\n", sink);
  return std::chrono::duration<double>(t1 - t0).count();
}

// shape functions / rules exposure for tests
int orc_shape(int kind, const double* x, double* N, double* dN) {
  const Basis& b = basis(kind);
  eval_basis(b, x, N, dN);
  return b.nnode;
}

int orc_gauss(int kind, int ngauss, double* xw) {  // returns npoint; xw: npoint × 4
  auto r = gauss_rule(kind, ngauss > 0 ? ngauss : default_ngauss(kind));
  for (size_t i = 0; i < r.size(); i++) { xw[i * 4] = r[i].x[0]; xw[i * 4 + 1] = r[i].x[1]; xw[i * 4 + 2] = r[i].x[2]; xw[i * 4 + 3] = r[i].w; }
  return (int)r.size();
}

// material point driver for unit tests: update then stiffness (ndim 2 or 3)
int orc_material_update(int ndim, int model, double E, double nu, double H, double z0, double* state /* ns+4 */,
                        const double* deps, double* D_out) {
  Param p; p.model = model; p.E = E; p.nu = nu; p.H = H; p.z0 = z0;
  p.K = E / (3.0 * (1.0 - 2.0 * nu)); p.G = E / (2.0 * (1.0 + nu));
  Ideal id; id.ndim = ndim;
  int ns = ndim == 2 ? 4 : 6;
  IPState s;
  for (int k = 0; k < ns; k++) s.sig[k] = state[k];
  s.z = state[ns]; s.lam = state[ns + 1]; s.elastic = state[ns + 2] != 0.0; s.epbar = state[ns + 3];
  int rc = 0;
  if (deps) { if (update_stress(p, id, ns, s, deps)) rc = 1; s.epbar += s.lam; }
  if (D_out) { if (stiffness(p, id, ns, s, D_out)) rc = 2; }
  for (int k = 0; k < ns; k++) state[k] = s.sig[k];
  state[ns] = s.z; state[ns + 1] = s.lam; state[ns + 2] = s.elastic ? 1.0 : 0.0; state[ns + 3] = s.epbar;
  return rc;
}

}  // extern "C"
