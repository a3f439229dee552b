// extern "C" entry points (include/pmsim_amd.h).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>

#include "pm_internal.hpp"

using namespace pm;

static thread_local std::string g_last_error;

namespace pm {
void set_global_error(const std::string& msg) { g_last_error = msg; }
}  // namespace pm

#define CHECK_HANDLE(f)                                                        \
  do {                                                                         \
    if (!(f)) { g_last_error = "null handle"; return 1; }                      \
  } while (0)
#define CHECK_INIT(f)                                                          \
  do {                                                                         \
    if (!(f)) { g_last_error = "null handle"; return 1; }                      \
    if (!(f)->initialized) { (f)->err = "pm_initialize has not been called"; return 1; } \
  } while (0)
#define CHECK_INIT_CONST(f)                                                    \
  do {                                                                         \
    if (!(f)) { g_last_error = "null handle"; return 1; }                      \
    if (!(f)->initialized) { const_cast<pm_fem*>(f)->err = "pm_initialize has not been called"; return 1; } \
  } while (0)

extern "C" {

const char* pm_last_error(const pm_fem* f) { return f ? f->err.c_str() : g_last_error.c_str(); }
const char* pm_global_last_error(void) { return g_last_error.c_str(); }

void pm_default_config(pm_config* c) {
  std::memset(c, 0, sizeof(pm_config));
  c->lagrange_mult_method = 0;
  c->linear_problem = 0;
  c->constant_tangent = 0;
  c->n_max_time_steps = 1000;
  c->n_max_iterations = 10;
  c->allowed_step_n_failure = 100;
  c->t_ini = 0.0; c->t_fin = 1.0; c->dt = 1.0; c->dt_out = 1.0;
  c->tol_rr_abs = 1e-10;
  c->tol_mdu_rel = 1e-8;
  c->symmetry_check_tolerance = 1e-10;
  c->lin_sol = 0;
  c->tangent_precision = 0;
  c->pcg_rtol = 1e-12;
  c->pcg_max_iter = 20000;
  c->keep_history = 0;
  c->tol_rr_scaled = 0.0;
  c->tangent_kernel = 0;
  c->use_graphs = 1;
  c->pcg_small_max = -1;
  c->arc_length_method = 0;
  c->arc_length_psi = 1.0;
  c->first_trial_loading_factor = 1.0;
}

int pm_fem_create(pm_fem** out, int32_t ndim, int64_t npoint, const double* coords, int64_t ncell, const int32_t* kinds,
                  const int32_t* attributes, const int64_t* cell_ptr, const int32_t* cell_points) {
  if (!out) { g_last_error = "null output handle"; return 1; }
  *out = nullptr;
  if (ndim != 2 && ndim != 3) { g_last_error = "ndim must be 2 or 3"; return 1; }
  if (npoint < 1) { g_last_error = "there must be at least one point in the mesh"; return 1; }
  if (ncell < 1) { g_last_error = "there are no cells in the mesh"; return 1; }
  if (!coords || !kinds || !attributes || !cell_ptr || !cell_points) { g_last_error = "null mesh array"; return 1; }
  if (cell_ptr[0] != 0) { g_last_error = "cell pointers must start at 0 and not decrease"; return 1; }
  for (int64_t e = 0; e < ncell; e++)
    if (cell_ptr[e + 1] < cell_ptr[e]) { g_last_error = "cell pointers must start at 0 and not decrease"; return 1; }
  auto* f = new pm_fem();
  f->ndim = ndim; f->npoint = npoint; f->ncell = ncell;
  f->X.assign(coords, coords + npoint * ndim);
  f->kind.assign(kinds, kinds + ncell);
  f->attr.assign(attributes, attributes + ncell);
  f->cptr.assign(cell_ptr, cell_ptr + ncell + 1);
  f->cpts.assign(cell_points, cell_points + cell_ptr[ncell]);
  pm_default_config(&f->cfg);
  *out = f;
  return 0;
}

void pm_fem_destroy(pm_fem* f) {
  if (!f) return;
  if (f->initialized) { (void)hipSetDevice(f->device); device_free(f); }
  delete f;
}

int pm_set_param_solid(pm_fem* f, int32_t attribute, int32_t model, double young, double poisson, double hh, double z_ini,
                       int32_t ngauss) {
  CHECK_HANDLE(f);
  if (model != 0 && model != 1) { f->err = "unknown stress-strain model"; return 1; }
  ParamSolid p;
  p.model = model; p.young = young; p.poisson = poisson; p.hh = hh; p.z_ini = z_ini; p.ngauss = ngauss; p.set = true;
  f->params[attribute] = p;
  return 0;
}

int pm_set_model_settings(pm_fem* f, int32_t attribute, int32_t general_plasticity, int32_t gp_interp_nn_max, double nle_beta) {
  CHECK_HANDLE(f);
  if (f->initialized) { f->err = "model settings must be set before pm_initialize"; return 1; }
  ModelSettings& m = f->model_settings[attribute];
  m.general_plasticity = general_plasticity != 0; m.gp_interp_nn_max = gp_interp_nn_max; m.nle_beta = nle_beta;
  return 0;
}

int pm_set_idealization(pm_fem* f, int32_t axisymmetric, int32_t plane_stress, double thickness) {
  CHECK_HANDLE(f);
  if (thickness <= 0.0) { f->err = "thickness must be > 0.0"; return 1; }
  if ((axisymmetric || plane_stress) && f->ndim != 2) { f->err = "axisymmetric or plane-stress requires ndim = 2"; return 1; }
  f->axisym = axisymmetric != 0; f->plane_stress = plane_stress != 0; f->thickness = thickness;
  return 0;
}

int pm_set_config(pm_fem* f, const pm_config* c) {
  CHECK_HANDLE(f);
  if (!c) { f->err = "null configuration"; return 1; }
  if (c->dt <= 0.0) { f->err = "Δt is smaller than the allowed minimum"; return 1; }
  if (c->tangent_precision < 0 || c->tangent_precision > 2) { f->err = "unknown tangent precision"; return 1; }
  if (f->initialized && c->tangent_precision != f->cfg.tangent_precision) { f->err = "tangent precision must be set before pm_initialize"; return 1; }
  if (c->tangent_kernel < 0 || c->tangent_kernel > 5) { f->err = "unknown tangent kernel"; return 1; }
  const bool retune = f->initialized && c->tangent_kernel != f->cfg.tangent_kernel;
  if (f->initialized) release_graphs(f);  // captured launches depend on the configuration
  f->cfg = *c;
  f->pcg_small_max = c->pcg_small_max >= 0 ? c->pcg_small_max : kPcgSmallMaxDefault;
  if (retune) return tune_tangent(f);
  return 0;
}

int pm_set_essential(pm_fem* f, int64_t n, const int32_t* points, const int32_t* dofs, const double* table, int32_t nstation) {
  CHECK_HANDLE(f);
  if (f->initialized) { f->err = "essential conditions must be set before pm_initialize"; return 1; }
  if (n < 0 || (n > 0 && (nstation < 1 || !points || !dofs || !table))) { f->err = "an essential table needs at least one station"; return 1; }
  f->ess_point.assign(points, points + n);
  f->ess_dof.assign(dofs, dofs + n);
  f->ess_table.assign(table, table + n * (int64_t)nstation);
  f->nstation = nstation;
  return 0;
}

int pm_set_point_loads(pm_fem* f, int64_t n, const int32_t* points, const int32_t* dofs, const double* table, int32_t nstation) {
  CHECK_HANDLE(f);
  if (n < 0 || (n > 0 && (!points || !dofs || !table))) { f->err = "null load table"; return 1; }
  if (f->initialized) { f->err = "natural conditions must be set before pm_initialize"; return 1; }
  if (nstation < 1) { f->err = "a load table needs at least one station"; return 1; }
  for (int64_t i = 0; i < n; i++)
    if (points[i] < 0 || points[i] >= f->npoint || dofs[i] < 0 || dofs[i] >= f->ndim) { f->err = "cannot find (PointId, Dof) in the DOF map"; return 1; }
  f->pl_point.assign(points, points + n);
  f->pl_dof.assign(dofs, dofs + n);
  f->pl_table.assign(table, table + n * (int64_t)nstation);
  f->pl_nstation = nstation;
  if (nstation > 1) f->loads_vary = true;
  return 0;
}

int pm_add_facet_qn(pm_fem* f, int32_t facet_kind, int32_t npts, const int32_t* points, int64_t cell, const double* values,
                    int32_t nstation) {
  return pm_add_facet_load(f, 0, facet_kind, npts, points, cell, values, nstation);
}

int pm_add_facet_load(pm_fem* f, int32_t nbc, int32_t facet_kind, int32_t npts, const int32_t* points, int64_t cell,
                      const double* values, int32_t nstation) {
  CHECK_HANDLE(f);
  if (!points || !values) { f->err = "null facet points or values"; return 1; }
  if (f->initialized) { f->err = "natural conditions must be set before pm_initialize"; return 1; }
  if (nbc < 0 || nbc > 3) { f->err = "unknown distributed load (0 Qn, 1 Qx, 2 Qy, 3 Qz)"; return 1; }
  if (kind_nnode(facet_kind) != npts) { f->err = "wrong number of points for facet kind"; return 1; }
  if (nstation < 1) { f->err = "a load table needs at least one station"; return 1; }
  const bool edge = kind_ndim(facet_kind) == 1;
  if (f->ndim == 2 && !edge) { f->err = "a 2D mesh takes edge facets only"; return 1; }
  if (nbc == 0 && f->ndim == 3 && edge) { f->err = "Qn natural boundary condition is not available for 3D edge"; return 1; }
  if (nbc == 3 && f->ndim == 2) { f->err = "cannot find the number of a (PointId, DOF) pair"; return 1; }
  if (nbc == 0 && (cell < 0 || cell >= f->ncell)) { f->err = "facet owner cell out of range"; return 1; }
  if (cell >= f->ncell) { f->err = "facet owner cell out of range"; return 1; }
  for (int32_t i = 0; i < npts; i++)
    if (points[i] < 0 || points[i] >= f->npoint) { f->err = "cannot find the number of a (PointId, DOF) pair"; return 1; }
  Facet F{facet_kind, std::vector<int32_t>(points, points + npts), cell < 0 ? -1 : cell, std::vector<double>(values, values + nstation)};
  F.nbc = nbc;
  f->facets.push_back(std::move(F));
  if (nstation > 1) f->loads_vary = true;
  return 0;
}

int pm_gauss_rule(int32_t kind, int32_t ngauss, int32_t* npoint, double* xi, double* w) {
  Rule R;
  if (kind < 0 || kind >= NKIND || !gauss_rule(kind, ngauss > 0 ? ngauss : kind_default_ngauss(kind), R)) {
    g_last_error = "no integration rule for this kind and number of points";
    return 1;
  }
  *npoint = R.n();
  if (xi) std::copy(R.x.begin(), R.x.end(), xi);
  if (w) std::copy(R.w.begin(), R.w.end(), w);
  return 0;
}

int pm_shape_values(int32_t kind, int32_t npoint, const double* xi, double* N) {
  const int nn = kind_nnode(kind);
  if (nn <= 0) { g_last_error = "unknown kind"; return 1; }
  for (int32_t p = 0; p < npoint; p++) shape(kind, xi + 3 * p, N + (size_t)p * nn, nullptr);
  return 0;
}

int pm_extrapolation_matrix(int32_t kind, int32_t ngauss, double* E) {
  std::vector<double> e;
  if (!extrapolation_matrix(kind, ngauss, e)) { g_last_error = "cannot compute the extrapolation matrix"; return 1; }
  std::copy(e.begin(), e.end(), E);
  return 0;
}

int pm_initialize(pm_fem* f, int32_t device) {
  if (!f) { g_last_error = "null handle"; return 1; }
  if (f->initialized || f->released) { f->err = "already initialized"; return 1; }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) { f->err = "no HIP device available"; return 1; }
  if (device < 0 || device >= ndev) { f->err = "invalid HIP device"; return 1; }
  f->device = device;
  if (build_groups(f) || build_patterns(f)) return 1;
  assemble_f_ext(f, 0);
  if (device_alloc(f)) { device_free(f); return 1; }
  noalloc_setup(f->device);  // the no-allocation entry point's scratch address on this device
  f->initialized = true;
  return tune_tangent(f);
}

int pm_initialize_internal_values(pm_fem* f) {
  CHECK_INIT(f);
  (void)hipSetDevice(f->device);
  for (auto& g : f->groups) launch_init_state(f, g);
  return hipStreamSynchronize(f->stream) == hipSuccess ? 0 : 1;
}

int pm_reset_algorithmic_variables(pm_fem* f) {
  CHECK_INIT(f);
  for (auto& g : f->groups) launch_reset_lambda(f, g);
  return hipStreamSynchronize(f->stream) == hipSuccess ? 0 : 1;
}

int pm_backup_secondary_values(pm_fem* f) {
  CHECK_INIT(f);
  for (auto& g : f->groups) launch_copy(f, g.d_backup, g.d_state, g.nel * g.nip * (f->nsig + 4));
  return hipStreamSynchronize(f->stream) == hipSuccess ? 0 : 1;
}

int pm_restore_secondary_values(pm_fem* f) {
  CHECK_INIT(f);
  for (auto& g : f->groups) launch_copy(f, g.d_state, g.d_backup, g.nel * g.nip * (f->nsig + 4));
  return hipStreamSynchronize(f->stream) == hipSuccess ? 0 : 1;
}

int pm_update_secondary_values(pm_fem* f) {
  CHECK_INIT(f);
  if (element_pass(f, true, false, false)) return 1;
  return check_device_error(f);
}

int pm_assemble_f_int(pm_fem* f) {
  CHECK_INIT(f);
  if (element_pass(f, false, true, false)) return 1;
  return check_device_error(f);
}

int pm_assemble_kk(pm_fem* f) {
  CHECK_INIT(f);
  if (gp_refuse(f, false, true)) return 1;
  for (auto& g : f->groups) launch_tangent(f, g);
  launch_scatter(f);
  return check_device_error(f);
}

int pm_element_pass(pm_fem* f) {
  CHECK_INIT(f);
  if (element_pass(f, true, true, true)) return 1;
  return check_device_error(f);
}

int pm_solve(pm_fem* f, int32_t* n_converged) {
  CHECK_INIT(f);
  if (!n_converged) { f->err = "null output"; return 1; }
  (void)hipSetDevice(f->device);
  return newton_solve(f, n_converged);
}

int pm_get_stats(const pm_fem* f, pm_stats* s) {
  CHECK_HANDLE(f);
  if (!s) return 1;
  *s = f->stats;
  s->n_ip = f->nip;
  s->n_block_nnz = f->nnzb;
  s->device_bytes = f->device_bytes;
  return 0;
}

int pm_sizes(const pm_fem* f, int64_t* out) {
  CHECK_HANDLE(f);
  if (!out) return 1;
  if (!f->initialized && !f->released) return 1;
  out[0] = f->ndof; out[1] = f->nprescribed; out[2] = f->nip; out[3] = f->nsig; out[4] = f->nnzb; out[5] = f->ndim;
  return 0;
}

int pm_set_ddu(pm_fem* f, const double* ddu) {
  CHECK_INIT(f);
  if (!ddu) { f->err = "null increment"; return 1; }
  PM_HIP(hipMemcpyAsync(f->d_ddu, ddu, sizeof(double) * f->ndof, hipMemcpyHostToDevice, f->stream));
  PM_HIP(hipStreamSynchronize(f->stream));
  return 0;
}

int pm_get_u(const pm_fem* f, double* u, double* lambda) {
  CHECK_HANDLE(f);
  if (!u) return 1;
  if (f->released) {
    std::copy(f->rel_u.begin(), f->rel_u.end(), u);
    if (lambda) std::copy(f->rel_lam.begin(), f->rel_lam.end(), lambda);
    return 0;
  }
  if (!f->initialized) return 1;
  if (hipMemcpyAsync(u, f->d_u, sizeof(double) * f->ndof, hipMemcpyDeviceToHost, f->stream) != hipSuccess) return 1;
  if (lambda && f->nprescribed &&
      hipMemcpyAsync(lambda, f->d_lam, sizeof(double) * f->nprescribed, hipMemcpyDeviceToHost, f->stream) != hipSuccess) return 1;
  return hipStreamSynchronize(f->stream) == hipSuccess ? 0 : 1;
}

int pm_get_f_int(const pm_fem* f, double* out) {
  CHECK_HANDLE(f);
  if (!out) return 1;
  if (!f->initialized) return 1;
  if (hipMemcpyAsync(out, f->d_fint, sizeof(double) * f->ndof, hipMemcpyDeviceToHost, f->stream) != hipSuccess) return 1;
  return hipStreamSynchronize(f->stream) == hipSuccess ? 0 : 1;
}

int pm_assemble_f_ext(pm_fem* f, double t) {
  CHECK_INIT(f);
  const double k = (t - f->cfg.t_ini) / f->cfg.dt;
  if (!std::isfinite(k)) { f->err = "Found NaN or Inf"; return 1; }
  assemble_f_ext(f, (int64_t)std::llround(k));
  PM_HIP(hipMemcpyAsync(f->d_fext, f->f_ext.data(), sizeof(double) * f->ndof, hipMemcpyHostToDevice, f->stream));
  PM_HIP(hipStreamSynchronize(f->stream));
  return 0;
}

int pm_get_f_ext(const pm_fem* f, double* out) {
  CHECK_INIT_CONST(f);
  if (!out) return 1;
  if (hipMemcpyAsync(out, f->d_fext, sizeof(double) * f->ndof, hipMemcpyDeviceToHost, f->stream) != hipSuccess) return 1;
  return hipStreamSynchronize(f->stream) == hipSuccess ? 0 : 1;
}

int pm_calculate_residuals_vector(pm_fem* f, double loading_factor) {
  CHECK_INIT(f);
  dev_residual_plain(f, loading_factor);
  PM_HIP(hipStreamSynchronize(f->stream));
  return 0;
}

int pm_get_residual(const pm_fem* f, double* out) {
  CHECK_INIT_CONST(f);
  if (!out) return 1;
  if (hipMemcpyAsync(out, f->d_rr, sizeof(double) * f->ndof, hipMemcpyDeviceToHost, f->stream) != hipSuccess) return 1;
  return hipStreamSynchronize(f->stream) == hipSuccess ? 0 : 1;
}

int pm_apply_kk(pm_fem* f, const double* x, double* y) {
  CHECK_INIT(f);
  if (!x || !y) { f->err = "null argument"; return 1; }
  double* dx = f->d_pcg;             // PCG work vectors: not live outside a linear solve
  double* dy = f->d_pcg + f->ndof;
  PM_HIP(hipMemcpyAsync(dx, x, sizeof(double) * f->ndof, hipMemcpyHostToDevice, f->stream));
  dev_spmv(f, dx, dy, false);
  PM_HIP(hipMemcpyAsync(y, dy, sizeof(double) * f->ndof, hipMemcpyDeviceToHost, f->stream));
  PM_HIP(hipStreamSynchronize(f->stream));
  return 0;
}

int pm_get_element_kk(const pm_fem* f, int64_t cell, double* kk) {
  CHECK_INIT_CONST(f);
  if (!kk) return 1;
  if (cell < 0 || cell >= f->ncell) { const_cast<pm_fem*>(f)->err = "cell out of range"; return 1; }
  const Group& g = f->groups[f->cell_group[cell]];
  if (g.general_plasticity) { const_cast<pm_fem*>(f)->err = "TODO"; return 1; }  // no stiffness (elastoplastic.rs:528-536)
  const int64_t el = f->cell_local[cell];
  const int nn = g.nnode, nd = g.ndim, ne = nn * nd;
  std::vector<double> v(g.kew);
  if (f->d_ke) {
    if (hipMemcpyAsync(v.data(), f->d_ke + g.ke_base + el * g.kew, sizeof(double) * g.kew, hipMemcpyDeviceToHost, f->stream) != hipSuccess) return 1;
  } else {
    std::vector<float> w(g.kew);
    if (hipMemcpyAsync(w.data(), f->d_ke32 + g.ke_base + el * g.kew, sizeof(float) * g.kew, hipMemcpyDeviceToHost, f->stream) != hipSuccess) return 1;
    if (hipStreamSynchronize(f->stream) != hipSuccess) return 1;
    std::copy(w.begin(), w.end(), v.begin());
  }
  if (hipStreamSynchronize(f->stream) != hipSuccess) return 1;
  int rk[32];
  cell_ranks(f, cell, rk);
  // stored: upper node blocks in rank order; block (a, b) of a lower rank pair is the transpose of (b, a)
  for (int a = 0; a < nn; a++)
    for (int b = 0; b < nn; b++) {
      const bool up = rk[a] <= rk[b];
      const double* blk = v.data() + (size_t)ke_block_offset(nn, nd, up ? rk[a] : rk[b], up ? rk[b] : rk[a]);
      for (int i = 0; i < nd; i++)
        for (int j = 0; j < nd; j++) kk[(size_t)(a * nd + i) * ne + b * nd + j] = up ? blk[i * nd + j] : blk[j * nd + i];
    }
  return 0;
}

int pm_get_element_f_int(const pm_fem* f, int64_t cell, double* out) {
  CHECK_INIT_CONST(f);
  if (!out) return 1;
  if (cell < 0 || cell >= f->ncell) { const_cast<pm_fem*>(f)->err = "cell out of range"; return 1; }
  const Group& g = f->groups[f->cell_group[cell]];
  const int64_t n = (int64_t)g.nnode * g.ndim;
  if (hipMemcpyAsync(out, f->d_fe + g.fe_base + f->cell_local[cell] * n, sizeof(double) * n, hipMemcpyDeviceToHost, f->stream) != hipSuccess)
    return 1;
  return hipStreamSynchronize(f->stream) == hipSuccess ? 0 : 1;
}

int pm_last_error_location(const pm_fem* f, int64_t* cell, int32_t* gauss) {
  CHECK_HANDLE(f);
  if (cell) *cell = f->err_cell;
  if (gauss) *gauss = f->err_gauss;
  return 0;
}

int pm_get_ip_state(const pm_fem* f, double* v) {
  CHECK_HANDLE(f);
  if (!v) return 1;
  if (f->released) { std::copy(f->rel_ip.begin(), f->rel_ip.end(), v); return 0; }
  if (!f->initialized) return 1;
  return copy_ip_state_to_host(f, v);
}

int pm_set_ip_state(pm_fem* f, const double* v) {
  CHECK_INIT(f);
  if (!v) { f->err = "null IP state"; return 1; }
  return copy_ip_state_to_device(f, v);
}

int pm_get_csr(const pm_fem* f, int64_t* rowptr, int32_t* cols, double* vals) {
  CHECK_HANDLE(f);
  if (!rowptr) return 1;
  if (!f->initialized) return 1;
  const int nd = f->ndim;
  std::vector<double> kv;
  if (vals && copy_kval_to_host(f, kv)) return 1;
  int64_t k = 0;
  rowptr[0] = 0;
  for (int64_t a = 0; a < f->npoint; a++)
    for (int i = 0; i < nd; i++) {
      for (int32_t blk = f->h_browptr[a]; blk < f->h_browptr[a + 1]; blk++)
        for (int j = 0; j < nd; j++) {
          if (cols) cols[k] = nd * f->h_bcols[blk] + j;
          if (vals) vals[k] = kv[(size_t)blk * nd * nd + i * nd + j];
          k++;
        }
      rowptr[a * nd + i + 1] = k;
    }
  return 0;
}

int pm_history_count(const pm_fem* f, int32_t* n) {
  CHECK_HANDLE(f);
  if (!n) return 1;
  *n = (int32_t)f->hist_u.size();
  return 0;
}

int pm_get_history(const pm_fem* f, int32_t index, double* u, double* lambda, double* ip_state, double* t) {
  CHECK_HANDLE(f);
  if (index < 0 || index >= (int32_t)f->hist_u.size()) return 1;
  if (u) std::memcpy(u, f->hist_u[index].data(), sizeof(double) * f->ndof);
  if (lambda && f->nprescribed) std::memcpy(lambda, f->hist_lam[index].data(), sizeof(double) * f->nprescribed);
  if (ip_state) std::memcpy(ip_state, f->hist_ip[index].data(), sizeof(double) * f->hist_ip[index].size());
  if (t) *t = f->hist_t[index];
  return 0;
}

int pm_get_history_loading_factor(const pm_fem* f, int32_t index, double* ell) {
  CHECK_HANDLE(f);
  if (index < 0 || index >= (int32_t)f->hist_ell.size() || !ell) return 1;
  *ell = f->hist_ell[index];
  return 0;
}

int pm_time_element_pass(pm_fem* f, int32_t reps, double* ms) {
  CHECK_INIT(f);
  if (!ms || reps < 1) { f->err = "pm_time_element_pass needs reps >= 1 and an output array"; return 1; }
  if (gp_refuse(f, true, true)) return 1;
  hipEvent_t ev[5];
  for (auto& e : ev) PM_HIP(hipEventCreate(&e));
  double acc[4] = {0, 0, 0, 0};
  float tot = 0.0f;
  for (int r = 0; r < reps; r++) {
    PM_HIP(hipEventRecord(ev[0], f->stream));
    for (auto& g : f->groups) launch_update(f, g, true);
    PM_HIP(hipEventRecord(ev[1], f->stream));
    launch_residual_gather(f);
    PM_HIP(hipEventRecord(ev[2], f->stream));
    for (auto& g : f->groups) launch_tangent(f, g);
    PM_HIP(hipEventRecord(ev[3], f->stream));
    launch_scatter(f);
    PM_HIP(hipEventRecord(ev[4], f->stream));
    PM_HIP(hipEventSynchronize(ev[4]));
    for (int k = 0; k < 4; k++) {
      float m = 0.0f;
      PM_HIP(hipEventElapsedTime(&m, ev[k], ev[k + 1]));
      acc[k] += m;
    }
    float m = 0.0f;
    PM_HIP(hipEventElapsedTime(&m, ev[0], ev[4]));
    tot += m;
  }
  for (int k = 0; k < 4; k++) ms[k] = acc[k] / reps;
  ms[4] = tot / reps;
  // the production path: the same pass replayed as one captured graph (or, without graphs, back to back)
  PM_HIP(hipEventRecord(ev[0], f->stream));
  for (int r = 0; r < reps; r++)
    if (element_pass(f, true, true, true)) { for (auto& e : ev) (void)hipEventDestroy(e); return 1; }
  PM_HIP(hipEventRecord(ev[1], f->stream));
  PM_HIP(hipEventSynchronize(ev[1]));
  float m = 0.0f;
  PM_HIP(hipEventElapsedTime(&m, ev[0], ev[1]));
  ms[5] = m / reps;
  for (auto& e : ev) (void)hipEventDestroy(e);
  f->stats.ms_update = ms[0]; f->stats.ms_residual = ms[1]; f->stats.ms_tangent = ms[2]; f->stats.ms_scatter = ms[3];
  return check_device_error(f);
}

int pm_linear_solve(pm_fem* f, const double* b, double* x, int32_t* iterations) {
  CHECK_INIT(f);
  if (!b || !x || !iterations) { f->err = "null argument"; return 1; }
  double* db = f->d_pcg + 6 * f->ndof;  // scratch beyond the PCG work vectors
  double* dx = f->d_mdu;
  PM_HIP(hipMemcpyAsync(db, b, sizeof(double) * f->ndof, hipMemcpyHostToDevice, f->stream));
  if (f->cfg.lin_sol == 0) build_block_jacobi(f);
  int it = 0;
  if (linear_solve_free(f, db, dx, &it)) return 1;
  *iterations = it;
  PM_HIP(hipMemcpyAsync(x, dx, sizeof(double) * f->ndof, hipMemcpyDeviceToHost, f->stream));
  PM_HIP(hipStreamSynchronize(f->stream));
  return 0;
}

}  // extern "C"
