// Internal structures of the MI355X solid element loop library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "../../include/pmsim_amd.h"

namespace pm {

enum Kind { TRI3 = 0, TRI6, TRI10, TRI15, QUA4, QUA8, QUA9, TET4, TET10, HEX8, HEX20, NKIND };
enum FacetKind { LIN2 = 100, LIN3, LIN4, LIN5 };

// Device error buffer (4 int32 words): word 0 holds the OR of the error bits raised by the element kernels,
// words 2-3 the smallest key (group << 40 | IP index within the group) of an IP that raised one, so the
// first failing cell and integration point can be reported as pmsim prints them (von_mises.rs:147-150).
enum DevErr { ERR_SIGD = 1, ERR_NORMS = 2, ERR_DETJ = 4, ERR_ASYM = 8 };
constexpr uint64_t kErrKeyNone = ~0ull;
constexpr int kErrKeyShift = 40;
__device__ inline void raise_error(int32_t* err, int bit, uint64_t key) {
  atomicOr(err, bit);
  atomicMin(reinterpret_cast<unsigned long long*>(err + 2), (unsigned long long)key);
}

int kind_nnode(int kind);
int kind_ndim(int kind);
int kind_default_ngauss(int kind);

// Integration rule: points (r,s,t) and weights
struct Rule {
  std::vector<double> x;  // npoint*3
  std::vector<double> w;  // npoint
  int n() const { return (int)w.size(); }
};
bool gauss_rule(int kind, int ngauss, Rule& r);
bool node_xi(int kind, int m, double* xi);
bool extrapolation_matrix(int kind, int ngauss, std::vector<double>& E);

// Shape functions N and derivatives dN/dxi (nnode*ndim) at a reference point (closed forms).
void shape(int kind, const double* x, double* N, double* dN);

struct ParamSolid {
  int model = 0;  // 0 linear elastic, 1 von Mises
  double young = 0, poisson = 0, hh = 0, z_ini = 0;
  int ngauss = 0;
  bool set = false;
};

// model settings of one cell attribute (src/material/settings.rs:6-50; Config::update_model_settings)
struct ModelSettings {
  bool general_plasticity = false;  // von Mises through the general elastoplastic formulation (elastoplastic.rs)
  int gp_interp_nn_max = 30;
  double nle_beta = 0.0;
};

// Per-kind element batch on the device.
struct PcgEngine;
struct PassEngine;

// Element matrices are stored as their upper node blocks in rank order (a <= b, the nodes of a cell sorted
// by global id), row after row, each block row-major; the lower blocks are the transposes.  Rank order makes
// the stored blocks exactly the upper CSR blocks they feed, and the blocks that consecutive CSR blocks of a
// row read adjacent in memory.  Block (a, b) starts at ke_block_offset(nn, nd, a, b).  Hex20 matrices carry
// one padding value after rank row 5, so that rows 0-5 and rows 6-19 (105 blocks each) are two halves of
// 946 values, each a whole number of 16-byte pieces in FP64.  Each element's storage is padded to a multiple
// of 4 values so that every element starts on a 16-byte boundary in FP32 and FP64.
constexpr int GATHER_CHUNK = 256;  // CSR blocks per gather workgroup (its block size)
// Persistent tangents hand out elements from kWorkCounters counters, kWorkStride int32 apart (one 256-byte
// line each): counter c owns elements c, c + n, c + 2n, ... (n counters in use), so that the device-scope
// atomics of the wavefronts are spread over as many addresses.
constexpr int kWorkCounters = 64, kWorkStride = 64;
constexpr int packed_block(int nn, int a, int b) { return a * (2 * nn - a + 1) / 2 + (b - a); }
constexpr int packed_ke_values(int nn, int nd) { return ((nn * (nn + 1) / 2 * nd * nd + 3) / 4) * 4; }
constexpr int ke_split_row(int nn, int nd) { return nn == 20 && nd == 3 ? 6 : nn; }
constexpr int ke_block_offset(int nn, int nd, int a, int b) {
  return packed_block(nn, a, b) * nd * nd + (a >= ke_split_row(nn, nd) ? 1 : 0);
}
static_assert(ke_block_offset(20, 3, 6, 6) == 946 && ke_block_offset(20, 3, 19, 19) + 9 == 1891 &&
              packed_ke_values(20, 3) == 1892, "Hex20 element-matrix halves");
// true when a group's element matrices, ew values each, lie inside the element-matrix buffer
inline bool ke_range_ok(int64_t ke_base, int64_t nel, int kew, int ew, int64_t n_ke_values) {
  return kew == ew && ke_base >= 0 && ke_base + nel * (int64_t)ew <= n_ke_values;
}

struct Group {
  int kind = 0, ndim = 2, nnode = 0, nip = 0, attribute = 0;
  ParamSolid param;
  bool general_plasticity = false;  // von Mises under the general formulation: its stiffness is not available
  // material constants
  double K = 0, G = 0;
  // plane-stress elastic (LinearElastic only)
  bool plane_stress = false;
  std::vector<int64_t> cells;      // global cell ids (ascending)
  int64_t nel = 0;
  int64_t ip_base = 0;             // first global IP index (cell order) of this group's first cell
  // device arrays
  int32_t* d_conn = nullptr;       // nel*nnode
  uint8_t* d_rank = nullptr;       // nel*nnode: rank of each local node among the cell's global ids
  double* d_tab = nullptr;         // [w (nip) | N (nip*nnode) | dN (nip*nnode*ndim)]
  double* d_geo = nullptr;         // [nel*nip][ndim*ndim + 2]: J^-1, det J and radius (axisymmetric) of each IP
  double* d_state = nullptr;       // SoA: [nsig+4][nel*nip]  (sigma..., z, lambda, elastic, epbar)
  double* d_backup = nullptr;
  int64_t fe_base = 0;             // offset (in doubles) of this group's f_e in the element force buffer
  int64_t ke_base = 0;             // offset (in values) of this group's element matrices
  int kew = 0;                     // values per element matrix: packed upper node blocks, padded (packed_ke_values)
  int tune_ch = 0;                 // IPs per chunk of the vector tangent chosen by the tuner (0: heuristic)
  int tune_kernel = 0;             // tangent kernel chosen at initialize: 1 node-block vector, 2 matrix core
};

// boundary facet with a distributed load: value at each time station (one entry: constant);
// nbc: 0 normal (Qn), 1/2/3 traction in x/y/z (Qx/Qy/Qz)
struct Facet { int kind; std::vector<int32_t> pts; int64_t cell; std::vector<double> values; int nbc = 0; };

struct Timers { double update = 0, residual = 0, tangent = 0, scatter = 0, linear = 0, total = 0; };

}  // namespace pm

// measured on MI355X: one workgroup wins below ~1.5k DOF (7-10 us per iteration), the graph-replayed
// multi-workgroup iteration (17-22 us, launch-bound up to 50k DOF) above
constexpr int64_t kPcgSmallMaxDefault = 1536;

struct pm_fem {
  // mesh
  int ndim = 2;
  int64_t npoint = 0, ncell = 0;
  std::vector<double> X;
  std::vector<int32_t> kind, attr;
  std::vector<int64_t> cptr;
  std::vector<int32_t> cpts;
  std::map<int, pm::ParamSolid> params;
  std::map<int, pm::ModelSettings> model_settings;
  bool axisym = false, plane_stress = false;
  double thickness = 1.0;
  pm_config cfg;
  // boundary conditions
  std::vector<int32_t> ess_point, ess_dof;
  std::vector<double> ess_table;
  int32_t nstation = 0;
  std::vector<int32_t> pl_point, pl_dof;
  std::vector<double> pl_table;        // point-load values at the time stations (n x pl_nstation)
  int32_t pl_nstation = 1;
  std::vector<pm::Facet> facets;
  bool loads_vary = false;             // some natural load changes with time: F_ext is rebuilt per step
  // derived sizes
  int nsig = 4;
  int64_t ndof = 0, nprescribed = 0, nip = 0;
  std::vector<int64_t> cell_ip;     // first IP of each cell (cell order)
  std::vector<int32_t> cell_group;  // group of each cell
  std::vector<int64_t> cell_local;  // index inside its group
  std::vector<pm::Group> groups;
  std::vector<int64_t> prescribed_eq;
  std::vector<uint8_t> is_prescribed;  // per DOF
  std::vector<double> f_ext;           // host-assembled external forces at the current station
  double ftf = 0.0;                    // F_ext . F_ext (serial, DOF order)
  // node-block CSR of K
  int bs = 2;                          // block size = ndim
  std::vector<int32_t> h_browptr;      // npoint+1  (full pattern: host direct solver, pm_get_csr)
  std::vector<int32_t> h_bcols;        // nnzb
  int64_t nnzb = 0;                    // node blocks of the full (symmetric) pattern
  // K is held on the device as its upper node blocks (b >= a); rows read their lower blocks through
  // the transposes of the stored mirrors (pm_spmv.hpp)
  int64_t nnzu = 0;                    // upper blocks = (nnzb + npoint) / 2
  std::vector<int32_t> h_urp, h_ucol;  // upper pattern (npoint+1, nnzu)
  std::vector<int32_t> h_lrp, h_lcol, h_lidx;  // lower blocks of each row: column, stored upper block
  std::vector<int32_t> h_full2up;      // full block -> stored upper block, sign bit: read transposed
  std::vector<int32_t> h_kmap_ptr, h_kmap;   // K gather map (host copy until upload)
  std::vector<int32_t> h_korder;             // gather chunks (TPB CSR blocks each) in processing order
  std::vector<int32_t> h_kwave;              // K gather by wavefronts: first CSR block of each wavefront's range
  std::vector<int32_t> h_fmap_ptr, h_fmap;   // R gather map
  // K gather by rows: per (node, cell) contribution in ascending cell order, the element's block row
  // (start block index, node count) and the CSR position of each of its blocks within the node's row
  // device
  int device = 0;
  int n_cu = 256;                      // compute units of the device
  int64_t lds_per_cu = 160 * 1024;     // LDS bytes per compute unit
  int64_t lds_per_block = 64 * 1024;   // LDS bytes one workgroup may allocate
  hipStream_t stream = nullptr;
  bool initialized = false;
  bool released = false;               // device buffers freed after a solve; results kept below
  std::vector<double> rel_u, rel_lam, rel_ip;
  double* d_X = nullptr;               // point coordinates
  int32_t* d_urp = nullptr;            // upper pattern
  int32_t* d_ucol = nullptr;
  int32_t* d_lrp = nullptr;            // lower blocks of each row
  int32_t* d_lcol = nullptr;
  int32_t* d_lidx = nullptr;
  int32_t* d_frp = nullptr;            // the full pattern and its map to the stored blocks, for the full K
  int32_t* d_fcol = nullptr;           // that the multi-workgroup PCG widens at each solve (full_pattern_alloc)
  int32_t* d_f2u = nullptr;
  double* d_kfull = nullptr;           // nnzb*bs*bs (FP64 tangents) or
  float* d_kfull32 = nullptr;          // nnzb*bs*bs (FP32 / BF16 tangents)
  double* d_kval = nullptr;            // nnzu*bs*bs (FP64), FP64 tangents only
  float* d_kval32 = nullptr;           // nnzu*bs*bs (FP32), reduced-precision tangents only
  double* d_ke = nullptr;              // element matrices FP64: upper node blocks in rank order (ke_block_offset)
  float* d_ke32 = nullptr;             // FP32 element blocks
  int64_t n_ke_values = 0;
  bool launch_refused = false;         // a launch whose output range did not fit its buffer was skipped
  const char* refuse_msg = nullptr;    // the refusal's message (null: an out-of-range element-matrix launch)
  int32_t* d_kmap_ptr = nullptr;       // nnzb+1
  int32_t* d_korder = nullptr;         // gather chunk processing order
  int32_t* d_kwave = nullptr;          // nkwave+1: whole CSR blocks of at most 64 contributions per wavefront
  int64_t nkwave = 0;
  int32_t* d_kmap = nullptr;           // contributions: element block index (ascending cell order)
  double* d_fe = nullptr;              // element forces
  int64_t n_fe = 0;
  int32_t* d_fmap_ptr = nullptr;       // npoint+1
  int32_t* d_fmap = nullptr;           // fe offsets (in units of ndim-vectors)
  double* d_u = nullptr;
  double* d_ddu = nullptr;
  double* d_fint = nullptr;
  double* d_fext = nullptr;
  double* d_rr = nullptr;              // residual (ndof)
  double* d_mdu = nullptr;             // correction (ndof)
  double* d_mdu0 = nullptr;            // first correction of the step (ndof + nprescribed)
  double* d_lam = nullptr;             // Lagrange multipliers (nprescribed)
  double* d_y = nullptr;               // multiplier corrections (nprescribed)
  double* d_cval = nullptr;            // prescribed values at the current time (nprescribed)
  int64_t* d_peq = nullptr;            // prescribed equations
  double* d_free = nullptr;            // 1 for free DOFs, 0 for prescribed (ndof)
  double* d_pcg = nullptr;             // PCG work vectors (6*ndof)
  double* d_arc = nullptr;             // arc-length vectors: u_old, u_anc, du1, F_free (ndof each), then
                                       // lambda_old, lambda_anc, dlambda, y1 (nprescribed each); on first use
  double* d_binv = nullptr;            // block-Jacobi inverses (npoint*bs*bs)
  double* d_red = nullptr;             // reduction partials
  double* d_scal = nullptr;            // device scalars
  int32_t* d_err = nullptr;            // device error flags and first failing IP (DevErr, raise_error)
  int32_t* d_work = nullptr;           // work counters of the persistent Hex20 tangents (reset before each launch)
  int64_t err_cell = -1;               // cell and Gauss point of the last device-side error (-1: none)
  int32_t err_gauss = -1;
  int64_t pcg_small_max = kPcgSmallMaxDefault;  // systems up to this many DOFs use the single-workgroup PCG
  pm::PcgEngine* engine = nullptr;     // batched PCG of a serving run (set by pm_serve), else per handle
  pm::PassEngine* pass_engine = nullptr;  // batched element passes of a serving run (set by the server)
  hipGraphExec_t pass_graph[8] = {};   // captured element passes, by (update, residual, tangent)
  hipGraphExec_t pcg_graph = nullptr;  // captured chunk of PCG iterations
  const double* pcg_graph_x = nullptr; // solution vector the chunk was captured for
  int64_t device_bytes = 0;
  // state
  double t = 0;
  double ell = 0;                      // loading factor (FemState::ell; 0 before the first step)
  std::vector<double> h_lam;           // host mirror of multipliers
  // history
  std::vector<std::vector<double>> hist_u, hist_lam, hist_ip;
  std::vector<double> hist_t, hist_ell;
  // stats
  pm_stats stats{};
  pm::Timers timers;
  std::string err;
};

namespace pm {

// host setup (pm_host.cpp)
int build_groups(pm_fem* f);
int build_patterns(pm_fem* f);
void assemble_f_ext(pm_fem* f, int64_t station);
int copy_kval_to_host(const pm_fem* f, std::vector<double>& kv);  // K block values in FP64 (widened from FP32)
// device (pm_kernels.hip)
int device_alloc(pm_fem* f);
void device_free(pm_fem* f);
void launch_update(pm_fem* f, Group& g, bool do_update);
void launch_residual_gather(pm_fem* f);
void launch_tangent(pm_fem* f, Group& g);
bool gp_refuse(pm_fem* f, bool do_update, bool tangent);
// matrix-core tangent (prec: 0 FP64, 1 FP32, 2 BF16 with FP32 accumulation); false if the kind has none
void cell_ranks(const pm_fem* f, int64_t e, int* rk);
bool launch_tangent_mfma(pm_fem* f, Group& g, int prec);
bool launch_tangent_hex20(pm_fem* f, Group& g);
bool tangent_hex20_fits(const pm_fem* f, const Group& g);
bool launch_tangent_hex20_mfma(pm_fem* f, Group& g);
bool tangent_hex20_mfma_fits(const pm_fem* f, const Group& g);
// choose the tangent kernel (and work split) of every group, timing the candidates when asked to
int tune_tangent(pm_fem* f);
// drops every captured graph (configuration, tuning or buffers changed)
void release_graphs(pm_fem* f);
// element passes of several handles in shared launches (serving): what each handle asks for
struct PassJob { pm_fem* f; bool update = false, do_update = false, residual = false, tangent = false; };
struct PassBatchBuf { void* d = nullptr; size_t cap = 0; std::vector<char> host; };  // staged descriptors and maps
// runs the jobs' passes on `stream` (returns after the stream is idle); 0 on success
int pass_batch_run(PassBatchBuf* buf, hipStream_t stream, std::vector<PassJob>& jobs);
PassEngine* pass_engine_create(int device);
void pass_engine_destroy(PassEngine* e);
// hands the handle's pass to the engine and waits for it (the handle's stream must be idle)
int pass_engine_run(PassEngine* e, const PassJob& job);
// the handle's element-pass pieces, through its pass engine when it has one, else on its own stream
int run_pass(pm_fem* f, bool update, bool do_update, bool residual, bool tangent);
// batched PCG engine shared by the requests of one serving run
PcgEngine* engine_create(int device);
void engine_destroy(PcgEngine* e);
int engine_solve(PcgEngine* e, pm_fem* f, const double* d_b, double* d_x, int* iters);
void launch_scatter(pm_fem* f);
void launch_init_state(pm_fem* f, Group& g);
void launch_geometry(pm_fem* f, Group& g);
void launch_reset_lambda(pm_fem* f, Group& g);
void launch_copy(pm_fem* f, double* dst, const double* src, int64_t n);
int check_device_error(pm_fem* f);
void set_global_error(const std::string& msg);  // pm_global_last_error's message (calls without a handle)
void reset_device_error(pm_fem* f);
size_t update_shared_bytes(const Group& g);
// vector kernels (pm_kernels.hip)
double dev_max_abs(pm_fem* f, const double* x, int64_t n, const double* mask);
double dev_max_rel(pm_fem* f, const double* x, const double* x0, int64_t n);
double dev_dot(pm_fem* f, const double* a, const double* b, int64_t n);
void dev_residual(pm_fem* f, double* g, double ell = 1.0);
void dev_residual_plain(pm_fem* f, double ell);  // R = F_int - ell F_ext only
void dev_rhs(pm_fem* f, const double* g, double* t, double* kt, double* b);
void dev_complete_correction(pm_fem* f, const double* g, double* x, double* kx, double* y, const double* rhs = nullptr);
void dev_arc_trial(pm_fem* f, int64_t n, bool extrapolate, double alpha, const double* u_old, const double* u_anc, double* u,
                   double* du);
void dev_axpy(pm_fem* f, int64_t n, double a, const double* x, double* y);
void dev_masked(pm_fem* f, const double* x, double* out);
void dev_update_primary(pm_fem* f, const double* x, const double* y);
void dev_fill(pm_fem* f, double* x, double v, int64_t n);
int arc_alloc(pm_fem* f);
int full_pattern_alloc(pm_fem* f);
// linear solvers
int pcg_solve(pm_fem* f, const double* d_b, double* d_x, int* iters);
// PCG over the stored upper-block K in caller-given work memory (work 4 ndof, red 3 x 2048, scal 4); allocates nothing
int pcg_solve_sym(pm_fem* f, const double* d_b, double* d_x, double* work, double* red, double* scal, int* iters);
// takes the device address of the no-allocation entry point's scratch array (pm_noalloc.hip) on `device`
// (current); called by pm_initialize
void noalloc_setup(int device);
int host_direct_solve(pm_fem* f, const std::vector<double>& b, std::vector<double>& x);
void build_block_jacobi(pm_fem* f);
void dev_spmv(pm_fem* f, const double* x, double* y, bool masked);
// Newton driver (pm_solver.cpp)
int newton_solve(pm_fem* f, int32_t* n_converged);
int element_pass(pm_fem* f, bool update, bool residual, bool tangent);
int linear_solve_free(pm_fem* f, const double* d_b, double* d_x, int* iters);
int copy_ip_state_to_host(const pm_fem* f, double* v);
int copy_ip_state_to_device(pm_fem* f, const double* v);

#define PM_HIP(call)                                                        \
  do {                                                                      \
    hipError_t e_ = (call);                                                 \
    if (e_ != hipSuccess) {                                                 \
      f->err = std::string("HIP error: ") + hipGetErrorString(e_) + " at " #call; \
      return 1;                                                             \
    }                                                                       \
  } while (0)

}  // namespace pm
