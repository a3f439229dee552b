// Shape functions (closed forms) and integration rules for the eleven solid kinds.
//
// Node orders follow gemlab's GeoKind conventions as used by pmsim's meshes
// (e.g. src/base/sample_meshes.rs:541-649 for Hex20, :305-360 for Tri15).  Derivatives
// dN/dxi are obtained by complex-step differentiation of the closed-form polynomials,
// which is exact to rounding for polynomials (no subtractive cancellation).
#include <cmath>
#include <complex>
#include <utility>

#include "pm_internal.hpp"

namespace pm {

int kind_nnode(int k) {
  static const int nn[NKIND] = {3, 6, 10, 15, 4, 8, 9, 4, 10, 8, 20};
  if (k >= 0 && k < NKIND) return nn[k];
  if (k >= LIN2 && k <= LIN5) return k - LIN2 + 2;
  return 0;
}
int kind_ndim(int k) { return (k >= 0 && k <= QUA9) ? 2 : (k < NKIND ? 3 : 1); }
// integration points when ParamSolid.ngauss is None (Tri3: 3, Qua8: 9 and Hex8: 8 are the counts
// pmsim writes in its own state files, data/results/artificial/*.json)
int kind_default_ngauss(int k) {
  static const int ng[NKIND] = {3, 3, 6, 12, 4, 9, 9, 1, 4, 8, 27};
  return (k >= 0 && k < NKIND) ? ng[k] : 0;
}

namespace {

template <class T>
T lagrange_lattice(int n, int p, T L) {  // prod_{m<n} (p L - m)/(m+1)
  T r = T(1.0);
  for (int m = 0; m < n; m++) r = r * (double(p) * L - double(m)) / double(m + 1);
  return r;
}

// barycentric lattice indices of simplex nodes
const int TRI_LAT[4][15][3] = {
    {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}},
    {{2, 0, 0}, {0, 2, 0}, {0, 0, 2}, {1, 1, 0}, {0, 1, 1}, {1, 0, 1}},
    // Tri10: 3 (1/3,0) 4 (2/3,1/3) 5 (0,2/3) 6 (2/3,0) 7 (1/3,2/3) 8 (0,1/3) 9 (1/3,1/3)
    {{3, 0, 0}, {0, 3, 0}, {0, 0, 3}, {2, 1, 0}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {0, 1, 2}, {2, 0, 1}, {1, 1, 1}},
    // Tri15 (r,s): 3 (.5,0) 4 (.5,.5) 5 (0,.5) 6 (.25,0) 7 (.75,0) 8 (.75,.25) 9 (.25,.75) 10 (0,.75)
    //              11 (0,.25) 12 (.25,.25) 13 (.5,.25) 14 (.25,.5);   L = (1-r-s, r, s) * 4
    {{4, 0, 0}, {0, 4, 0}, {0, 0, 4}, {2, 2, 0}, {0, 2, 2}, {2, 0, 2}, {3, 1, 0}, {1, 3, 0},
     {0, 3, 1}, {0, 1, 3}, {1, 0, 3}, {3, 0, 1}, {2, 1, 1}, {1, 2, 1}, {1, 1, 2}},
};

template <class T>
void values(int kind, const T* x, T* N) {
  const T r = x[0], s = x[1], t = x[2];
  switch (kind) {
    case TRI3: case TRI6: case TRI10: case TRI15: {
      int p = kind - TRI3 + 1, nn = kind_nnode(kind);
      T L[3] = {T(1.0) - r - s, r, s};
      for (int i = 0; i < nn; i++)
        N[i] = lagrange_lattice(TRI_LAT[p - 1][i][0], p, L[0]) * lagrange_lattice(TRI_LAT[p - 1][i][1], p, L[1]) *
               lagrange_lattice(TRI_LAT[p - 1][i][2], p, L[2]);
      return;
    }
    case QUA4: case QUA8: case QUA9: {
      static const double xi[9][2] = {{-1, -1}, {1, -1}, {1, 1}, {-1, 1}, {0, -1}, {1, 0}, {0, 1}, {-1, 0}, {0, 0}};
      if (kind == QUA4) {
        for (int i = 0; i < 4; i++) N[i] = 0.25 * (T(1.0) + xi[i][0] * r) * (T(1.0) + xi[i][1] * s);
      } else if (kind == QUA8) {
        for (int i = 0; i < 4; i++)
          N[i] = 0.25 * (T(1.0) + xi[i][0] * r) * (T(1.0) + xi[i][1] * s) * (xi[i][0] * r + xi[i][1] * s - 1.0);
        for (int i = 4; i < 8; i++) {
          if (xi[i][0] == 0.0) N[i] = 0.5 * (T(1.0) - r * r) * (T(1.0) + xi[i][1] * s);
          else N[i] = 0.5 * (T(1.0) + xi[i][0] * r) * (T(1.0) - s * s);
        }
      } else {
        auto l = [](double a, T z) -> T {  // quadratic Lagrange on {-1, 0, 1}
          if (a < 0) return 0.5 * z * (z - 1.0);
          if (a > 0) return 0.5 * z * (z + 1.0);
          return T(1.0) - z * z;
        };
        for (int i = 0; i < 9; i++) N[i] = l(xi[i][0], r) * l(xi[i][1], s);
      }
      return;
    }
    case TET4: case TET10: {
      T L[4] = {T(1.0) - r - s - t, r, s, t};
      if (kind == TET4) { for (int i = 0; i < 4; i++) N[i] = L[i]; return; }
      for (int i = 0; i < 4; i++) N[i] = L[i] * (2.0 * L[i] - 1.0);
      static const int e[6][2] = {{0, 1}, {1, 2}, {0, 2}, {0, 3}, {1, 3}, {2, 3}};
      for (int k = 0; k < 6; k++) N[4 + k] = 4.0 * L[e[k][0]] * L[e[k][1]];
      return;
    }
    case HEX8: case HEX20: {
      static const double xi[20][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1}, {-1, -1, 1}, {1, -1, 1}, {1, 1, 1},
                                       {-1, 1, 1}, {0, -1, -1}, {1, 0, -1}, {0, 1, -1}, {-1, 0, -1}, {0, -1, 1}, {1, 0, 1},
                                       {0, 1, 1}, {-1, 0, 1}, {-1, -1, 0}, {1, -1, 0}, {1, 1, 0}, {-1, 1, 0}};
      if (kind == HEX8) {
        for (int i = 0; i < 8; i++) N[i] = 0.125 * (T(1.0) + xi[i][0] * r) * (T(1.0) + xi[i][1] * s) * (T(1.0) + xi[i][2] * t);
        return;
      }
      for (int i = 0; i < 8; i++)
        N[i] = 0.125 * (T(1.0) + xi[i][0] * r) * (T(1.0) + xi[i][1] * s) * (T(1.0) + xi[i][2] * t) *
               (xi[i][0] * r + xi[i][1] * s + xi[i][2] * t - 2.0);
      for (int i = 8; i < 20; i++) {
        if (xi[i][0] == 0.0) N[i] = 0.25 * (T(1.0) - r * r) * (T(1.0) + xi[i][1] * s) * (T(1.0) + xi[i][2] * t);
        else if (xi[i][1] == 0.0) N[i] = 0.25 * (T(1.0) + xi[i][0] * r) * (T(1.0) - s * s) * (T(1.0) + xi[i][2] * t);
        else N[i] = 0.25 * (T(1.0) + xi[i][0] * r) * (T(1.0) + xi[i][1] * s) * (T(1.0) - t * t);
      }
      return;
    }
    case LIN2: case LIN3: case LIN4: case LIN5: {
      int n = kind - LIN2 + 2;
      double nodes[5];
      nodes[0] = -1.0; nodes[1] = 1.0;
      for (int i = 1; i < n - 1; i++) nodes[i + 1] = -1.0 + 2.0 * i / (n - 1);
      for (int i = 0; i < n; i++) {
        T v = T(1.0);
        for (int j = 0; j < n; j++) if (j != i) v = v * (r - nodes[j]) / (nodes[i] - nodes[j]);
        N[i] = v;
      }
      return;
    }
  }
}

}  // namespace

void shape(int kind, const double* x, double* N, double* dN) {
  int nn = kind_nnode(kind), nd = kind_ndim(kind);
  double xr[3] = {x[0], nd > 1 ? x[1] : 0.0, nd > 2 ? x[2] : 0.0};
  values<double>(kind, xr, N);
  if (!dN) return;
  const double h = 1e-30;
  typedef std::complex<double> C;
  for (int d = 0; d < nd; d++) {
    C xc[3] = {C(xr[0], 0), C(xr[1], 0), C(xr[2], 0)};
    xc[d] = C(xr[d], h);
    C Nc[20];
    values<C>(kind, xc, Nc);
    for (int m = 0; m < nn; m++) dN[m * nd + d] = Nc[m].imag() / h;
  }
}

// ----------------------------------------------------------------------------------------------
// integration rules (r fastest for tensor rules; simplex rules from Dunavant / Keast tables)
// ----------------------------------------------------------------------------------------------
static bool legendre(int n, double* x, double* w) {
  switch (n) {
    case 1: x[0] = 0; w[0] = 2; return true;
    case 2: x[0] = -1.0 / std::sqrt(3.0); x[1] = -x[0]; w[0] = w[1] = 1.0; return true;
    case 3: x[0] = -std::sqrt(0.6); x[1] = 0; x[2] = -x[0]; w[0] = w[2] = 5.0 / 9.0; w[1] = 8.0 / 9.0; return true;
    case 4: {
      double a = std::sqrt(3.0 / 7.0 - 2.0 / 7.0 * std::sqrt(1.2)), b = std::sqrt(3.0 / 7.0 + 2.0 / 7.0 * std::sqrt(1.2));
      double wa = (18.0 + std::sqrt(30.0)) / 36.0, wb = (18.0 - std::sqrt(30.0)) / 36.0;
      x[0] = -b; x[1] = -a; x[2] = a; x[3] = b; w[0] = wb; w[1] = wa; w[2] = wa; w[3] = wb;
      return true;
    }
  }
  return false;
}

bool gauss_rule(int kind, int ngauss, Rule& R) {
  R.x.clear(); R.w.clear();
  auto add = [&](double a, double b, double c, double w) { R.x.push_back(a); R.x.push_back(b); R.x.push_back(c); R.w.push_back(w); };
  int nd = kind_ndim(kind);
  bool tensor = (kind >= QUA4 && kind <= QUA9) || kind == HEX8 || kind == HEX20 || (kind >= LIN2 && kind <= LIN5);
  if (tensor) {
    int n1 = 0;
    for (int k = 1; k <= 4; k++) { int t = 1; for (int d = 0; d < nd; d++) t *= k; if (t == ngauss) n1 = k; }
    double x[4], w[4];
    if (!n1 || !legendre(n1, x, w)) return false;
    if (nd == 1) for (int i = 0; i < n1; i++) add(x[i], 0, 0, w[i]);
    else if (nd == 2) for (int j = 0; j < n1; j++) for (int i = 0; i < n1; i++) add(x[i], x[j], 0, w[i] * w[j]);
    else for (int k = 0; k < n1; k++) for (int j = 0; j < n1; j++) for (int i = 0; i < n1; i++) add(x[i], x[j], x[k], w[i] * w[j] * w[k]);
    return true;
  }
  if (kind >= TRI3 && kind <= TRI15) {
    auto p3 = [&](double a, double b, double w) { add(a, a, 0, w); add(b, a, 0, w); add(a, b, 0, w); };
    auto p6 = [&](double a, double b, double c, double w) {
      add(a, b, 0, w); add(b, a, 0, w); add(b, c, 0, w); add(c, b, 0, w); add(c, a, 0, w); add(a, c, 0, w);
    };
    switch (ngauss) {
      case 1: add(1.0 / 3.0, 1.0 / 3.0, 0, 0.5); return true;
      case 3: add(1.0 / 6.0, 1.0 / 6.0, 0, 1.0 / 6.0); add(2.0 / 3.0, 1.0 / 6.0, 0, 1.0 / 6.0); add(1.0 / 6.0, 2.0 / 3.0, 0, 1.0 / 6.0); return true;
      case 6:
        p3(0.445948490915965, 0.10810301816807, 0.5 * 0.223381589678011);
        p3(0.091576213509771, 0.816847572980459, 0.5 * 0.109951743655322);
        return true;
      case 7:
        add(1.0 / 3.0, 1.0 / 3.0, 0, 0.5 * 0.225);
        p3(0.470142064105115, 0.059715871789770, 0.5 * 0.132394152788506);
        p3(0.101286507323456, 0.797426985353087, 0.5 * 0.125939180544827);
        return true;
      case 12:
        p3(0.249286745170910, 0.501426509658179, 0.5 * 0.116786275726379);
        p3(0.063089014491502, 0.873821971016996, 0.5 * 0.050844906370207);
        p6(0.053145049844817, 0.310352451033784, 0.636502499121399, 0.5 * 0.082851075618374);
        return true;
    }
    return false;
  }
  if (kind == TET4 || kind == TET10) {
    switch (ngauss) {
      case 1: add(0.25, 0.25, 0.25, 1.0 / 6.0); return true;
      case 4: {
        double a = (5.0 - std::sqrt(5.0)) / 20.0, b = (5.0 + 3.0 * std::sqrt(5.0)) / 20.0;
        add(a, a, a, 1.0 / 24); add(b, a, a, 1.0 / 24); add(a, b, a, 1.0 / 24); add(a, a, b, 1.0 / 24);
        return true;
      }
      case 5:
        add(0.25, 0.25, 0.25, -2.0 / 15.0);
        add(1.0 / 6, 1.0 / 6, 1.0 / 6, 3.0 / 40); add(0.5, 1.0 / 6, 1.0 / 6, 3.0 / 40);
        add(1.0 / 6, 0.5, 1.0 / 6, 3.0 / 40); add(1.0 / 6, 1.0 / 6, 0.5, 3.0 / 40);
        return true;
    }
    return false;
  }
  return false;
}

// natural coordinates of node m of a solid kind (the nodes at which N_m = 1)
bool node_xi(int kind, int m, double* xi) {
  xi[0] = xi[1] = xi[2] = 0.0;
  const int nn = kind_nnode(kind);
  if (m < 0 || m >= nn) return false;
  switch (kind) {
    case TRI3: case TRI6: case TRI10: case TRI15: {
      const int p = kind - TRI3 + 1;
      xi[0] = double(TRI_LAT[p - 1][m][1]) / p;
      xi[1] = double(TRI_LAT[p - 1][m][2]) / p;
      return true;
    }
    case QUA4: case QUA8: case QUA9: {
      static const double q[9][2] = {{-1, -1}, {1, -1}, {1, 1}, {-1, 1}, {0, -1}, {1, 0}, {0, 1}, {-1, 0}, {0, 0}};
      xi[0] = q[m][0]; xi[1] = q[m][1];
      return true;
    }
    case TET4: case TET10: {
      static const double v[4][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
      static const int e[6][2] = {{0, 1}, {1, 2}, {0, 2}, {0, 3}, {1, 3}, {2, 3}};
      if (m < 4) { for (int d = 0; d < 3; d++) xi[d] = v[m][d]; return true; }
      for (int d = 0; d < 3; d++) xi[d] = 0.5 * (v[e[m - 4][0]][d] + v[e[m - 4][1]][d]);
      return true;
    }
    case HEX8: case HEX20: {
      static const double h[20][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1}, {-1, -1, 1}, {1, -1, 1}, {1, 1, 1},
                                      {-1, 1, 1}, {0, -1, -1}, {1, 0, -1}, {0, 1, -1}, {-1, 0, -1}, {0, -1, 1}, {1, 0, 1},
                                      {0, 1, 1}, {-1, 0, 1}, {-1, -1, 0}, {1, -1, 0}, {1, 1, 0}, {-1, 1, 0}};
      for (int d = 0; d < 3; d++) xi[d] = h[m][d];
      return true;
    }
  }
  return false;
}

namespace {

// inverse of a small symmetric positive definite matrix (n <= 27) by Gauss-Jordan with partial pivoting
bool small_inverse(std::vector<double> a, int n, std::vector<double>& inv) {
  inv.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; i++) inv[(size_t)i * n + i] = 1.0;
  for (int k = 0; k < n; k++) {
    int p = k;
    for (int i = k + 1; i < n; i++) if (std::fabs(a[(size_t)i * n + k]) > std::fabs(a[(size_t)p * n + k])) p = i;
    if (std::fabs(a[(size_t)p * n + k]) < 1e-300) return false;
    if (p != k)
      for (int j = 0; j < n; j++) { std::swap(a[(size_t)k * n + j], a[(size_t)p * n + j]); std::swap(inv[(size_t)k * n + j], inv[(size_t)p * n + j]); }
    const double d = a[(size_t)k * n + k];
    for (int j = 0; j < n; j++) { a[(size_t)k * n + j] /= d; inv[(size_t)k * n + j] /= d; }
    for (int i = 0; i < n; i++) {
      if (i == k) continue;
      const double f = a[(size_t)i * n + k];
      if (f == 0.0) continue;
      for (int j = 0; j < n; j++) { a[(size_t)i * n + j] -= f * a[(size_t)k * n + j]; inv[(size_t)i * n + j] -= f * inv[(size_t)k * n + j]; }
    }
  }
  return true;
}

// Moore-Penrose pseudo-inverse of a full-rank r x c matrix A (c x r result): (A^T A)^-1 A^T when
// r >= c, A^T (A A^T)^-1 otherwise
bool pseudo_inverse(const std::vector<double>& A, int r, int c, std::vector<double>& P) {
  P.assign((size_t)c * r, 0.0);
  std::vector<double> G, Gi;
  if (r >= c) {
    G.assign((size_t)c * c, 0.0);
    for (int i = 0; i < c; i++) for (int j = 0; j < c; j++) for (int k = 0; k < r; k++) G[(size_t)i * c + j] += A[(size_t)k * c + i] * A[(size_t)k * c + j];
    if (!small_inverse(G, c, Gi)) return false;
    for (int i = 0; i < c; i++) for (int j = 0; j < r; j++) for (int k = 0; k < c; k++) P[(size_t)i * r + j] += Gi[(size_t)i * c + k] * A[(size_t)j * c + k];
  } else {
    G.assign((size_t)r * r, 0.0);
    for (int i = 0; i < r; i++) for (int j = 0; j < r; j++) for (int k = 0; k < c; k++) G[(size_t)i * r + j] += A[(size_t)i * c + k] * A[(size_t)j * c + k];
    if (!small_inverse(G, r, Gi)) return false;
    for (int i = 0; i < c; i++) for (int j = 0; j < r; j++) for (int k = 0; k < r; k++) P[(size_t)i * r + j] += A[(size_t)k * c + i] * Gi[(size_t)k * r + j];
  }
  return true;
}

}  // namespace

// Extrapolation from the integration points to the nodes (pmsim's PostProc::nodal_stress,
// src/fem/post_processing.rs:455-520, through gemlab's get_extrap_matrix): nodal = E (nnode x nip)
// times the IP values.  With N (nip x nnode) the shape functions at the IPs:
//  * nip >= nnode: E = N^+ = (N^T N)^-1 N^T, the least-squares fit (N^-1 when square);
//  * ndim+1 <= nip < nnode: Durand R, Farias MM (2014) "A local extrapolation method for finite
//    elements", Adv. Eng. Software 67:1-9: E = N^+ (I - X X^+) + Xn X^+ with X (nip x ndim+1) the IPs'
//    natural coordinates and a column of ones and Xn (nnode x ndim+1) the same for the nodes, so that a
//    field linear in the natural coordinates is reproduced exactly;
//  * fewer IPs: E = N^+ (minimum norm; reproduces constant fields).
// The first case agrees with any least-squares extrapolation; the second is the published method
// named above.
bool extrapolation_matrix(int kind, int ngauss, std::vector<double>& E) {
  const int nn = kind_nnode(kind), nd = kind_ndim(kind);
  if (nn <= 0 || kind >= NKIND) return false;
  Rule R;
  if (!gauss_rule(kind, ngauss > 0 ? ngauss : kind_default_ngauss(kind), R)) return false;
  const int np = R.n();
  std::vector<double> N((size_t)np * nn);
  for (int p = 0; p < np; p++) shape(kind, &R.x[3 * p], &N[(size_t)p * nn], nullptr);
  std::vector<double> Np;
  if (!pseudo_inverse(N, np, nn, Np)) return false;  // nn x np
  if (np >= nn || np < nd + 1) { E = Np; return true; }
  const int c = nd + 1;
  std::vector<double> X((size_t)np * c), Xn((size_t)nn * c), Xp;
  for (int p = 0; p < np; p++) { for (int d = 0; d < nd; d++) X[(size_t)p * c + d] = R.x[3 * p + d]; X[(size_t)p * c + nd] = 1.0; }
  for (int m = 0; m < nn; m++) {
    double xi[3];
    node_xi(kind, m, xi);
    for (int d = 0; d < nd; d++) Xn[(size_t)m * c + d] = xi[d];
    Xn[(size_t)m * c + nd] = 1.0;
  }
  if (!pseudo_inverse(X, np, c, Xp)) return false;  // c x np
  // P = I - X X^+ (np x np)
  std::vector<double> P((size_t)np * np, 0.0);
  for (int i = 0; i < np; i++) {
    P[(size_t)i * np + i] = 1.0;
    for (int j = 0; j < np; j++) for (int k = 0; k < c; k++) P[(size_t)i * np + j] -= X[(size_t)i * c + k] * Xp[(size_t)k * np + j];
  }
  E.assign((size_t)nn * np, 0.0);
  for (int m = 0; m < nn; m++)
    for (int j = 0; j < np; j++) {
      double s = 0.0;
      for (int i = 0; i < np; i++) s += Np[(size_t)m * np + i] * P[(size_t)i * np + j];
      for (int k = 0; k < c; k++) s += Xn[(size_t)m * c + k] * Xp[(size_t)k * np + j];
      E[(size_t)m * np + j] = s;
    }
  return true;
}

}  // namespace pm
