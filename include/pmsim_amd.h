/*
 * pmsim_amd.h -- C ABI of the MI355X-native implicit solid element loop.
 *
 * One opaque handle (pm_fem) holds what pmsim spreads over FemBase + FemState +
 * SolverImplicit for the solid / von Mises path: the mesh grouped into per-kind element
 * batches, the device-resident integration-point state, the CSR (node-block) tangent and
 * the Newton driver.  Every function returns 0 on success; on failure it returns a
 * non-zero status and pm_last_error() gives the message, using pmsim's own error texts
 * where pmsim has one (src/lib.rs:2 StrError).
 *
 * Threading: a handle is not thread-safe; it owns one HIP stream (default priority) and
 * every call is ordered on it.  Different handles may be used from different threads.
 *
 * Kind codes: Tri3=0 Tri6=1 Tri10=2 Tri15=3 Qua4=4 Qua8=5 Qua9=6 Tet4=7 Tet10=8 Hex8=9
 * Hex20=10; boundary facets add Lin2=100 Lin3=101 Lin4=102 Lin5=103 (2D edges) and use
 * Tri3/Tri6/Qua4/Qua8 codes for 3D faces.  Local node orders follow gemlab's GeoKind.
 * DOF numbering is pmsim's closed form for solids: eq = ndim*point + dof
 * (src/base/all_dofs.rs:87-117, element_dofs.rs:123-131).
 */
#ifndef PMSIM_AMD_H
#define PMSIM_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pm_fem pm_fem;

/* Solver knobs.  Field meaning and defaults follow pmsim's Config (src/base/config.rs:184-231);
 * fields after `symmetry_check_tolerance` select the MI355X linear solver and tangent precision. */
typedef struct pm_config {
  int32_t lagrange_mult_method;     /* config.rs:426  (default 0) */
  int32_t linear_problem;           /* config.rs:371  (default 0) */
  int32_t constant_tangent;         /* config.rs:389  (default 0) */
  int32_t n_max_time_steps;         /* default 1000 */
  int32_t n_max_iterations;         /* default 10 */
  int32_t allowed_step_n_failure;   /* default 100 */
  double t_ini, t_fin, dt, dt_out;  /* defaults 0, 1, 1, 1 (constant Δt) */
  double tol_rr_abs;                /* default 1e-10 */
  double tol_mdu_rel;               /* default 1e-8 */
  double symmetry_check_tolerance;  /* < 0 disables (default 1e-10) */
  int32_t lin_sol;                  /* 0: device FP64 block-Jacobi PCG, 1: host direct (Cholesky) */
  int32_t tangent_precision;        /* 0: FP64, 1: FP32, 2: BF16 inputs with FP32 accumulation */
  double pcg_rtol;                  /* PCG stop: ||r|| <= pcg_rtol * ||b||  (default 1e-12) */
  int32_t pcg_max_iter;             /* default 20000 */
  int32_t keep_history;             /* store u and IP state of every output step on the host */
  double tol_rr_scaled;             /* > 0 adds ||R||inf <= tol_rr_scaled * ||F_int||inf as a convergence test */
  int32_t tangent_kernel;           /* 0: default (vector kernel for FP64 / FP32, matrix cores for BF16), 1: node-block vector kernel (FP64 arithmetic;
                                       persistent form for Hex20 with 27 IPs, same terms, agrees to rounding), 2: matrix-core (MFMA) kernel in the tangent precision
                                       (FP32 / BF16: persistent form for Hex20 with 27 IPs and Poisson ratio >= 0),
                                       3: tune at pm_initialize (fastest measured; FP64 variants agree to rounding, not bitwise),
                                       4: node-block vector kernel with one workgroup per group of elements (never the persistent form),
                                       5: matrix-core kernel per group of elements (never the persistent Hex20 form) */
  int32_t use_graphs;               /* 1 (default): element passes and 16-iteration PCG chunks replay captured HIP graphs */
  int32_t pcg_small_max;            /* systems with at most this many DOFs run the whole PCG in one workgroup; < 0: default */
  int32_t verbose;                  /* to stderr; bit 0: a line per converged step, bit 1: a line per Newton iteration
                                       (pmsim's verbose_timesteps / verbose_iterations, config.rs:166-173; default 0) */
  int32_t arc_length_method;        /* config.rs:395 (default 0): path following by the arc-length method
                                       (src/fem/arc_length_control.rs); the loads are scaled by the loading factor */
  double arc_length_psi;            /* config.rs:406 (default 1; 0 <= psi <= 1): weight of the load term in the constraint */
  double first_trial_loading_factor;/* config.rs:411 (default 1; |l0| >= 1e-12): loading factor of the first step */
} pm_config;

/* Statistics of the last pm_solve / element pass. */
typedef struct pm_stats {
  int32_t n_steps;                  /* converged time steps */
  int32_t n_iterations;             /* Newton iterations (all steps) */
  int64_t n_pcg_iterations;         /* PCG iterations (all solves) */
  int32_t n_host_fallbacks;         /* linear solves that fell back to the host direct solver */
  double ms_update, ms_residual, ms_tangent, ms_scatter, ms_linear, ms_total;
  int64_t n_ip;                     /* integration points */
  int64_t n_block_nnz;              /* node-block nonzeros of K */
  int64_t device_bytes;             /* device memory held by the handle */
} pm_stats;

/* ---- errors ---------------------------------------------------------------------------- */
const char* pm_last_error(const pm_fem* fem);      /* message of the last failure on `fem` */
const char* pm_global_last_error(void);            /* message of the last pm_fem_create failure */
/* Cell and integration point of the first IP that raised the last device-side material or Jacobian error
 * (pmsim prints them as "cell_id = .., gauss_id = .." before returning the error, von_mises.rs:147-150,
 * 213-216; this library prints the same line on stdout); -1 when the last error had none. */
int pm_last_error_location(const pm_fem* fem, int64_t* cell, int32_t* gauss);
void pm_default_config(pm_config* cfg);            /* fills pmsim's defaults (config.rs:184-231) */

/* ---- problem definition (replaces gemlab Mesh + FemBase::new, src/fem/fem_base.rs:25-30) -- */
int pm_fem_create(pm_fem** out, int32_t ndim, int64_t npoint, const double* coords, int64_t ncell,
                  const int32_t* kinds, const int32_t* attributes, const int64_t* cell_ptr,
                  const int32_t* cell_points);
void pm_fem_destroy(pm_fem* fem);

/* ParamSolid + StressStrain (src/base/parameters.rs:6-68,276-285); model 0 LinearElastic, 1 VonMises;
 * ngauss <= 0 selects the kind's default rule. */
int pm_set_param_solid(pm_fem* fem, int32_t attribute, int32_t model, double young, double poisson,
                       double hh, double z_ini, int32_t ngauss);
/* Model settings of one attribute (Config::update_model_settings, src/base/config.rs:515-517; Settings,
 * src/material/settings.rs:6-66), validated when the elements are allocated (pm_initialize) with pmsim's
 * texts.  general_plasticity selects, for von Mises, the general Elastoplastic model
 * (model_stress_strain.rs:65-71), whose stiffness pmsim does not provide (elastoplastic.rs:528-536): any
 * tangent then fails with pmsim's "TODO"; the update without a tangent is refused as not available. */
int pm_set_model_settings(pm_fem* fem, int32_t attribute, int32_t general_plasticity, int32_t gp_interp_nn_max,
                          double nle_beta);
/* Idealization (src/base/idealization.rs): 2D plane strain unless axisymmetric or plane stress. */
int pm_set_idealization(pm_fem* fem, int32_t axisymmetric, int32_t plane_stress, double thickness);
int pm_set_config(pm_fem* fem, const pm_config* cfg);
/* Essential (src/base/essential.rs:38-135): n prescribed (point, dof) with values at the time
 * stations t_k = t_ini + k*dt, table is n x nstation row-major (value = c * m(t_k)). */
int pm_set_essential(pm_fem* fem, int64_t n, const int32_t* points, const int32_t* dofs,
                     const double* table, int32_t nstation);
/* Natural point loads (src/fem/bc_concentrated.rs:46-55; Natural::points / points_fn,
 * src/base/natural.rs:92-123): table is n x nstation row-major with the values at the time stations
 * t_k (nstation = 1: constant in time). */
int pm_set_point_loads(pm_fem* fem, int64_t n, const int32_t* points, const int32_t* dofs, const double* table,
                       int32_t nstation);
/* Natural normal distributed load Qn on one boundary facet (src/fem/bc_distributed.rs:166-174;
 * Natural::edge(s)/face(s) and their _fn forms, src/base/natural.rs:56-140): values at the time
 * stations (nstation = 1: constant). */
int pm_add_facet_qn(pm_fem* fem, int32_t facet_kind, int32_t npts, const int32_t* points, int64_t cell,
                    const double* values, int32_t nstation);
/* Distributed load of kind nbc on one boundary facet (BcDistributed::calc_f_ext, src/fem/bc_distributed.rs:156-236):
 * nbc 0 = Qn (normal; the owner cell orients the outward normal; not on 3D edges), 1/2/3 = Qx/Qy/Qz (a traction
 * in x/y/z per unit facet measure; the owner cell is not used and may be -1; allowed on 3D edges).  Errors use
 * pmsim's texts: "Qn natural boundary condition is not available for 3D edge", and "cannot find the number of
 * a (PointId, DOF) pair" for Qz in 2D. */
int pm_add_facet_load(pm_fem* fem, int32_t nbc, int32_t facet_kind, int32_t npts, const int32_t* points, int64_t cell,
                      const double* values, int32_t nstation);

/* Builds element batches, the node-block CSR pattern, the deterministic scatter maps and all
 * device buffers (replaces SolverImplicit::new / SolverData::new / LinearSystem::new,
 * src/fem/solver_implicit.rs:27-51, solver_data.rs:37-82, linear_system.rs:89-181). */
int pm_initialize(pm_fem* fem, int32_t device);

/* ---- batched ElementTrait over device-resident state (src/fem/element_trait.rs:1-38) ------ */
int pm_initialize_internal_values(pm_fem* fem);  /* elements.rs -> von_mises.rs:107-116 */
int pm_reset_algorithmic_variables(pm_fem* fem); /* element_solid.rs:241-246 */
int pm_backup_secondary_values(pm_fem* fem);     /* element_solid.rs:227-231 */
int pm_restore_secondary_values(pm_fem* fem);    /* element_solid.rs:234-238 */
/* element_solid.rs:190-224: per IP strain increment from the device ddu, then the stress update
 * starting from the backup state (the driver's backup/restore protocol, solver_implicit.rs:254-263). */
int pm_update_secondary_values(pm_fem* fem);
int pm_assemble_f_int(pm_fem* fem);              /* element_solid.rs:124-138 + assembly.rs:38-46 */
int pm_assemble_kk(pm_fem* fem);                 /* element_solid.rs:168-185 + assembly.rs:70-139 */
/* One element pass = update + f_int + K (the benchmarked unit). */
int pm_element_pass(pm_fem* fem);

/* ---- Newton driver (SolverImplicit::solve, src/fem/solver_implicit.rs:56-272) ---------- */
int pm_solve(pm_fem* fem, int32_t* n_converged);
int pm_get_stats(const pm_fem* fem, pm_stats* stats);

/* ---- integration points and post-processing (no handle, no device) ---------------------- */
/* Integration rule of a solid kind (gemlab's Gauss::new_or_sized as called at
 * src/fem/element_solid.rs:59-62; ngauss <= 0: the kind's default): npoint, natural coordinates
 * xi (npoint x 3, unused trailing components 0) and weights w; xi / w may be NULL to query npoint. */
int pm_gauss_rule(int32_t kind, int32_t ngauss, int32_t* npoint, double* xi, double* w);
/* Shape functions N (npoint x nnode) of a kind (solid or facet) at natural coordinates xi
 * (npoint x 3): gemlab's interpolation as used by PostProc::gauss_coords (src/fem/post_processing.rs:147-160). */
int pm_shape_values(int32_t kind, int32_t npoint, const double* xi, double* N);
/* Extrapolation matrix E (nnode x nip, row-major) from the integration points to the nodes, as used
 * by PostProc::nodal_stress / nodal_stresses (src/fem/post_processing.rs:455-520, 637-653):
 * least-squares fit when nip >= nnode, Durand & Farias (2014) local extrapolation when
 * ndim+1 <= nip < nnode, minimum-norm fit below that. */
int pm_extrapolation_matrix(int32_t kind, int32_t ngauss, double* E);

/* ---- data movement -------------------------------------------------------------------- */
int pm_sizes(const pm_fem* fem, int64_t* out /* [ndof, n_prescribed, nip, nsig, n_block_nnz, ndim] */);
int pm_set_ddu(pm_fem* fem, const double* ddu /* ndof */);
int pm_get_u(const pm_fem* fem, double* u /* ndof */, double* lambda /* n_prescribed or NULL */);
int pm_get_f_int(const pm_fem* fem, double* f /* ndof */);
/* External forces at time t (SolverData::assemble_ff_int_and_ff_ext's F_ext half, src/fem/solver_data.rs:85-108:
 * distributed loads bc_distributed.rs:156-236 and point loads bc_concentrated.rs:46-55 at the time station
 * nearest t, prescribed equations skipped under RSM), assembled on the host and copied to the device. */
int pm_assemble_f_ext(pm_fem* fem, double t);
int pm_get_f_ext(const pm_fem* fem, double* f /* ndof */);
/* R = F_int - loading_factor * F_ext on the device (SolverData::calculate_residuals_vector, solver_data.rs:111-120),
 * from the last F_int and F_ext; the Lagrange-multiplier rows are not added (bc_prescribed.rs:99-108 does that). */
int pm_calculate_residuals_vector(pm_fem* fem, double loading_factor);
int pm_get_residual(const pm_fem* fem, double* r /* ndof */);
/* y = K x over all DOFs with the last assembled tangent (the global matrix of LinearSystem, linear_system.rs:65-70,
 * applied as an operator; FP32-held K is read in FP32 and accumulated in FP64). */
int pm_apply_kk(pm_fem* fem, const double* x /* ndof */, double* y /* ndof */);
/* Local tangent of one cell from the last pm_assemble_kk (ElementTrait::calc_jacobian, element_trait.rs:23;
 * ElementSolid::calc_jacobian, element_solid.rs:168-185): (nnode*ndim)^2 values, row-major, local DOF order
 * ndim*m + d (element_dofs.rs:123-131). */
int pm_get_element_kk(const pm_fem* fem, int64_t cell, double* kk);
/* Local internal force of one cell from the last update (ElementTrait::calc_f_int, element_solid.rs:124-138):
 * nnode*ndim values in local DOF order. */
int pm_get_element_f_int(const pm_fem* fem, int64_t cell, double* f);
/* IP state in cell order: per IP [sigma (nsig Mandel), z, lambda, elastic(0/1), epbar]. */
int pm_get_ip_state(const pm_fem* fem, double* v);
int pm_set_ip_state(pm_fem* fem, const double* v);
/* Global K (ndof x ndof) in scalar CSR: rowptr (ndof+1), cols (nnz), vals (nnz) as FP64 numbers (for
 * FP32 / BF16 tangents K is held in FP32 and its values are returned widened). */
int pm_get_csr(const pm_fem* fem, int64_t* rowptr, int32_t* cols, double* vals);
/* History of output steps kept with keep_history: count, then u (n x ndof) and IP states. */
int pm_history_count(const pm_fem* fem, int32_t* n);
int pm_get_history(const pm_fem* fem, int32_t index, double* u, double* lambda, double* ip_state, double* t);
/* Loading factor of an output step (FemState::ell, src/fem/fem_state.rs:17): 0 for the initial state, 1 after
 * a step without the arc-length method, the path's loading factor with it. */
int pm_get_history_loading_factor(const pm_fem* fem, int32_t index, double* ell);
/* Times `reps` element passes with HIP events on the handle's stream: per kernel class with events
 * between the launches, then `reps` passes as the solver runs them (one captured graph per pass when
 * use_graphs is set), in ms per pass. */
int pm_time_element_pass(pm_fem* fem, int32_t reps, double* ms /* [update, residual, tangent, scatter, total, pass] */);
/* Linear solve of the current K with a given right-hand side on the free DOFs (testing and
 * benchmarking the PCG): x = K_ff^-1 b_f, prescribed entries of x are zero. */
int pm_linear_solve(pm_fem* fem, const double* b, double* x, int32_t* iterations);
/* The same solve for callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc); it has no counterpart in pmsim.  b and x are device pointers of
 * ndof doubles owned by the caller.  Allocates nothing: the PCG work vectors, partial sums and scalars live
 * in one fixed 64 MiB device array of the library (taken per device at pm_initialize), so systems are
 * limited to 2,000,000 DOFs (larger ones are refused with an error); every call on a device shares that
 * array, so calls that overlap on one device race.  Block-Jacobi PCG over the stored upper-block K with
 * eager launches (no graph, no full-pattern copy of K) and no host fallback: a solve that does not reach
 * pcg_rtol within pcg_max_iter returns "PCG did not converge" with x as it stands. */
int pm_linear_solve_noalloc(pm_fem* fem, const double* d_b, double* d_x, int32_t* iterations);

/* ---- serving: many independent solves on one GPU ------------------------------------------------
 * pmsim solves one problem per SolverImplicit::new + solve (src/fem/solver_implicit.rs:27-178); the
 * entry points below run many such problems (each a fully configured, not yet initialized handle)
 * concurrently on one device. */

/* Per-request record filled by pm_serve (times in seconds from the start of pm_serve). */
typedef struct pm_request_stats {
  double t_arrival, t_admit, t_done;
  int32_t status;                   /* 0 = solved; otherwise the failing call's code (pm_last_error(handle) has the text) */
  int32_t n_steps, n_iterations;    /* converged steps, Newton iterations */
  int64_t n_pcg_iterations;
  int64_t device_bytes;             /* footprint estimate used for admission */
} pm_request_stats;

/* Upper bound of the device memory this handle allocates: at pm_initialize (mesh, batches, state,
 * element matrices, K pattern bound, vectors) and later in pm_solve (arc-length vectors, the
 * full-pattern K copy of the multi-workgroup PCG). */
int pm_estimate_device_bytes(const pm_fem* fem, int64_t* bytes);
/* Copies u, lambda and the IP state to host memory and frees every device buffer of the handle; the
 * getters pm_get_u / pm_get_ip_state / pm_get_history / pm_get_stats keep working from the copies. */
int pm_release_device(pm_fem* fem);
/* Continuous batching of independent solves.  Request i arrives at arrival_s[i] (seconds after the
 * call); requests are admitted in arrival order while fewer than max_active run and the sum of their
 * footprint estimates stays within hbm_budget_bytes (a lone request is always admitted).  Each admitted
 * request is initialized on `device` with its own stream, solved (pm_solve) and released
 * (pm_release_device) on a host worker thread, so the element loops and PCG iterations of the running
 * requests execute concurrently.  Returns 0 when every request was processed (check out[i].status). */
int pm_serve(int32_t n, pm_fem* const* fems, const double* arrival_s, int32_t device, int64_t hbm_budget_bytes,
             int32_t max_active, pm_request_stats* out);

/* The same serving loop with open submission: requests are submitted while others run and polled for.
 * A request is a configured, not yet initialized handle; it is queued at pm_server_submit (t_arrival), admitted
 * in submission order while fewer than max_active run and the footprint estimates of the running ones fit
 * hbm_budget_bytes (a lone request always runs), then initialized, solved (pm_solve) and released on a worker
 * thread with its own stream.  The element passes (update, residual gather, tangent, K gather) of all running
 * requests are launched together, one launch per kernel class over the union of their element groups, and
 * so are their multi-workgroup PCG iterations; each request's results are bitwise those of its own pm_solve.
 * Times in the records are seconds since pm_server_create.  pm_server_destroy waits for every submitted
 * request; the handles stay the caller's (results through pm_get_u / pm_get_ip_state / pm_get_stats). */
typedef struct pm_server pm_server;
int pm_server_create(pm_server** out, int32_t device, int64_t hbm_budget_bytes, int32_t max_active);
int pm_server_submit(pm_server* server, pm_fem* fem, int64_t* ticket /* 0, 1, ... in submission order */);
/* 0 and the record once the request is done (check out->status), 1 while it is queued or running, -1 for an
 * unknown ticket. */
int pm_server_poll(pm_server* server, int64_t ticket, pm_request_stats* out);
int pm_server_wait(pm_server* server, int64_t ticket, pm_request_stats* out);
/* Batching so far: out[0] = element-pass batches launched, out[1] = request passes they carried. */
int pm_server_stats(pm_server* server, int64_t* out /* [2] */);
int pm_server_destroy(pm_server* server);

#ifdef __cplusplus
}
#endif

#endif /* PMSIM_AMD_H */
