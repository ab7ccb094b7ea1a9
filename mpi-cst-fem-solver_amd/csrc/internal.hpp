// internal.hpp -- shared state of the MI355X assemble-and-solve library.
//
// Device data layout (all resident in HBM for the life of a context):
//   mesh:      xy[2nn] f64, tri[3ne] i32, E_elem[ne] f64 (optional)
//   node->elements:  ne_ptr[nn+1], ne_list[3ne] (element ids ascending per node)
//   pattern:   block CSR of 2x2 node blocks: brow_ptr[nb+1], bcol[nblk] (ascending per row),
//              bdiag[nb] = position of the diagonal block; the reference's scalar CSR is the
//              2x expansion (rows 2n, 2n+1 share [2m, 2m+1 for m in adj(n)]).
//   slot map:  slot_ptr[nblk+1], slot_src[9ne]: contributions to each block sorted by
//              element, encoded e*9 + a*3 + b (a, b = local node of the row / column).
//   matrix:    SELL-64 of 2x2 blocks: slice s holds block rows 64s..64s+63; its w_s block
//              columns are stored column-major with stride 64, so lane l of a wavefront
//              reads block k of row 64s+l at SELL position  sell_off[s] + 64k + l  (4 bytes
//              of column index).  FP64 values are planar per (slice, k): the 64 row-0 pairs
//              (a00 a01, 1 KB) then the 64 row-1 pairs (a10 a11, 1 KB), so that each half
//              is one fully coalesced 16-byte-per-lane load (sell_vidx).  FP32 copies keep
//              the 4 values of a block together (16 bytes per lane, already coalesced).
//              Padding blocks are zero with the row's own column.  sell_pos[q] maps CSR
//              block q to its SELL position.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <exception>
#include <memory>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "cstfem_hip.h"

namespace cstfem_hip {

constexpr int kSlice = 64;  // block rows per SELL slice = one wavefront

// Index in the FP64 SELL value array of entry j (0 a00, 1 a01, 2 a10, 3 a11) of the 2x2 block
// at SELL position pos (slice offsets are multiples of 64, so pos & 63 is the lane).
__host__ __device__ inline int64_t sell_vidx(int64_t pos, int j) {
  return 4 * (pos & ~int64_t(63)) + ((j >> 1) << 7) + 2 * (pos & 63) + (j & 1);
}

// 3x3 row-major constitutive matrix, passed to kernels by value
struct Mat3 {
  double d[9];
};

// element.cpp:7-15, evaluated on the host with the same operations as the device code
inline Mat3 constitutive_host(double E, double nu, int strain) {
  Mat3 m{};
  if (!strain) {
    const double c = E / (1.0 - nu * nu);
    const double v[9] = {c, c * nu, 0.0, c * nu, c, 0.0, 0.0, 0.0, c * (1.0 - nu) / 2.0};
    for (int i = 0; i < 9; ++i) m.d[i] = v[i];
  } else {
    const double c = E / ((1.0 + nu) * (1.0 - 2.0 * nu));
    const double v[9] = {c * (1.0 - nu), c * nu, 0.0, c * nu, c * (1.0 - nu), 0.0,
                         0.0, 0.0, c * (1.0 - 2.0 * nu) / 2.0};
    for (int i = 0; i < 9; ++i) m.d[i] = v[i];
  }
  return m;
}

// Error carrying a C-ABI status code.
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void hip_check(hipError_t e, const char* what, const char* file, int line);
#define HIP_CHECK(x) ::cstfem_hip::hip_check((x), #x, __FILE__, __LINE__)

template <class T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  ~DBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  void alloc(size_t count) {
    if (count == n && p) return;
    release();
    if (count == 0) return;
    if (hipMalloc(&p, count * sizeof(T)) != hipSuccess)
      throw Error(CSTFEM_ERR_ALLOC, "hipMalloc of " + std::to_string(count * sizeof(T)) + " bytes failed");
    n = count;
  }
  T* get() const { return p; }
};

// Scalars of one PCG solve, resident on the device.
// Written only by workgroup 0 of a kernel, or by the host between kernels.
struct PcgScalars {
  double pAp;      // p'Ap of a breakdown
  double rr;       // ||b - A x||^2 of the last true-residual evaluation
  double bnorm, rel, rtol;
  double delta;    // FP32 refinement: refresh r when ||r|| drops by this factor
  double rel_ref;  // FP32 refinement: ||r||/||b|| at the last refresh
  double pad0;
  double best_rel[2];  // flexible CG: smallest ||r||/||b|| so far, per iteration parity
  int k_in[2];     // iterations completed, written by iterations of parity 0 / 1
  int status, max_iter, record, refine_mode;
  int best_k[2];   // flexible CG: iteration of best_rel, per parity
  int stall_window;
  int pad1;
  // flexible CG's rate test (slow_window > 0): every slow_window iterations the best
  // ||r||/||b|| must have dropped to slow_factor times its value at the previous check
  // (slow_rel, slow_k: that value and iteration, per parity)
  double slow_rel[2];
  int slow_k[2];
  int slow_window;
  int pad2;
  double slow_factor;
};

// kRefine: the FP32 iteration asks the host for an FP64 true-residual refresh.
// kStalled: flexible CG made no progress for stall_window iterations, or too little over
// slow_window iterations (rate test).
enum { kRunning = 0, kConverged = 1, kBreakdown = 2, kMaxIter = 3, kRefine = 4, kStalled = 5 };

constexpr int kMpnnRounds = 3;  // message-passing rounds of the learned preconditioner

// Learned preconditioner state (mpnn.hip): FP32 master parameters with Adam moments, BF16
// weight copies, per-node activations, the decoded blocks and M's SELL values, and the
// training pairs (r_k, e_k) as FP32 rows of length n.
struct Mpnn {
  bool initialized = false;
  bool built = false;  // gsell holds M for the current parameters and matrix
  int step = 0;        // Adam steps taken
  int npairs = 0;
  DBuf<float> theta, grad, am, av;
  DBuf<uint16_t> wt[kMpnnRounds + 1], wn[kMpnnRounds + 1];  // [out x in], [in x out]
  DBuf<uint16_t> feat;
  DBuf<uint16_t> H[kMpnnRounds + 1], X[kMpnnRounds + 1];
  DBuf<float> ghat, dghat, gsell;
  DBuf<int> btrans;
  DBuf<float> R, E;
  // workspace of a training step (loss and backward), grown on demand and kept
  DBuf<float> wsU, wsV, wsdH, wsdX, wsCP, wsP;
  DBuf<uint16_t> wsdZ;
  DBuf<double> wsPart, wsW;
};

// One level of the aggregation multilevel preconditioner (mg.hip).  Level 0 is the context's
// matrix (2x2-block SELL); levels >= 1 hold scalar SELL-64 matrices with 3 DOFs per node
// (the rigid-body modes of one aggregate of the finer level).
struct MgLevel {
  int n = 0;      // DOFs
  int nodes = 0;  // nodes = n / bs
  int bs = 2;
  // block pattern over the nodes (host; ascending columns), device copy for levels >= 1
  std::vector<int> bptr, bcol;
  DBuf<int> d_bptr, d_bcol;
  // scalar SELL-64 matrix (levels >= 1): rows contiguous, slice s = rows 64s..64s+63; a
  // row's entries come in triples of one 3x3 block (slice widths are multiples of 3), so the
  // column indices are stored once per block: entry k of a row in slice s has column
  // 3 scol[off[s] / 3 + (k / 3) * 64 + lane] + k % 3
  int nslices = 0;
  std::vector<int64_t> h_off;
  DBuf<int64_t> off;
  DBuf<int> scol;
  DBuf<double> val;
  DBuf<float> val32;  // FP32 copy of val for the cycle's row sums (when enabled)
  DBuf<double> dinv;
  double lam = 0.0;  // estimate of lambda_max(D^-1 A)
  // restriction / prolongation to level + 1 (absent on the coarsest level)
  int nagg = 0;
  std::vector<int> h_agg;
  DBuf<int> agg, mptr, mem;  // aggregate of every node; member nodes of every aggregate
  DBuf<double> P;            // n x 3: rows of the tentative prolongator
  // level 0 of the learned smoother's cycle (coarse_fp32): P rounded to FP32, which that
  // cycle's restriction and prolongation read (the Galerkin product reads P)
  DBuf<float> P32;
  bool p32 = false;
  DBuf<double> nns;          // n x 3 near-null space of levels >= 1 (R factors of the level above)
  // work vectors
  DBuf<double> rhs, x, x2, d, res, kx, kres, kres2, kp, kq, kz;
  // fused residual + restriction on this level: aggregates per workgroup and the most member
  // rows one workgroup holds (0: separate kernels)
  int rr_ag = 0, rr_cap = 0;
  DBuf<double> part;  // reduction partials of the K-cycle steps
  // before reuse by a new hierarchy: sizes and bounds cleared, and the buffer that only some
  // hierarchies fill (val32, read whenever present) released; the others are reallocated by
  // the set-up, which keeps a buffer of the same size as it is
  void reset() {
    n = nodes = 0;
    bs = 2;
    nslices = 0;
    lam = 0.0;
    nagg = 0;
  }
};

struct Multigrid {
  bool ready = false;
  int fine_smoother = 0;  // 0: learned M, 1: Chebyshev on D^-1 A
  int kdepth = 3;         // levels 1..kdepth use the K-cycle, deeper ones the V-cycle
  int kdepth_set = 3;     // the set-up's value while a Chebyshev fallback widens kdepth
  std::vector<MgLevel*> lv;
  // level objects of the dropped hierarchy, in level order: the next set-up takes them back,
  // so that a hierarchy of the same sizes (the same mesh with new values) allocates nothing
  std::vector<MgLevel*> spare;
  int nc = 0;             // DOFs of the coarsest level (dense inverse)
  int tail_from = -1;     // first level of the fused single-workgroup tail cycle (-1: none)
  double m_bound = 0.0;   // lambda_max(M A) of the learned level-0 smoother
  bool coarse_fp32 = false;  // levels >= 1 hold FP32 copies of their values (learned smoother)
  // fused cycle kernels (CSTFEM_MG_FUSE=0: separate kernels): on levels >= 1 of fewer than
  // fuse_n DOFs, residual + restriction in one kernel, and where their row sums also take 4 or
  // more lanes the prolongation inside the first post-smoothing step; a K-cycle update left to
  // the kernel that reads it next where that kernel's workgroups read at most fold_max
  // partials in all
  bool fuse = true;
  int fuse_n = 200000;
  int64_t fold_max = 200000;
  // the set-up's lane-per-aggregate kernels on levels of many aggregates (CSTFEM_MG_SETUP_LANES;
  // 0: the wavefront forms everywhere, same results)
  bool setup_lanes = true;
  DBuf<double> cinv;
  DBuf<int> inv_flag;  // the device inverse's "not positive definite" flag
  // operands of the packed one-level tail (k_tail1), gathered at set-up for level tail_packed
  // (-1: none; the cycle then runs k_tail)
  DBuf<double> tail_pd;
  DBuf<int> tail_pi;
  int tail_packed = -1;
  // device workspaces of the set-up, grown on demand and kept, so that repeated set-ups
  // allocate nothing: member-list scratch, coarse degrees and error flags (iwork), the
  // aggregates the lane forms of the set-up kernels hand to the wavefront forms (big: a count,
  // then the list), power-iteration vectors (dwork)
  DBuf<int> iwork, big;
  DBuf<double> dwork;
  DBuf<double> dwork2;             // the learned smoother's check, on the auxiliary stream
  DBuf<double> dwork3;             // the levels' smoothing bounds, on the auxiliary stream
  DBuf<double> lam_dev;            // those bounds on the device, one per level
  hipEvent_t bound_ev = nullptr;   // the check starts after this point of the main stream
  hipEvent_t lam_ev = nullptr;     // a level's bound starts after its matrix
  std::vector<hipEvent_t> lvl_ev;  // the same, one per level (the set-up's bound helper)
  int* pinned = nullptr;  // page-locked staging of the fine node graph, kept across set-ups
  size_t pinned_n = 0;
  // page-locked copy of the coarse block pattern just built (the next level's host
  // aggregation reads it once pat_ev has passed), and the event after each aggregate-map
  // upload from `pinned` (the next level overwrites the staging after it)
  int* pinned_pat = nullptr;
  size_t pinned_pat_n = 0;
  hipEvent_t pat_ev = nullptr;
  hipEvent_t map_ev = nullptr;
  // the prepared aggregation's column download in pieces (one event each): its root pass
  // starts once the row pointers are in and waits for each piece when it reaches it
  static constexpr int kPrePieces = 8;
  hipEvent_t pre_piece[kPrePieces] = {};
  // level-0 aggregation started ahead of the set-up (mg_prepare): a worker thread waits for
  // the node graph's download (auxiliary stream, event pre_done), runs the greedy pass into
  // pre_agg and stages the map for upload; the set-up joins it.  Uses `pinned` until joined.
  std::thread pre_worker;
  bool pre_started = false;
  bool auto_prepare = false;  // mg_prepare was called: pattern builds start the aggregation
  std::vector<int> pre_agg;
  int pre_nagg = 0;
  bool pre_staged = false;  // the worker also copied the aggregate map into `pinned`
  std::exception_ptr pre_error;
  hipEvent_t pre_queued = nullptr, pre_done = nullptr;
  void join_prepared() {
    if (pre_worker.joinable()) pre_worker.join();
    pre_started = false;
  }
  MgLevel* take_level() {
    MgLevel* L;
    if (!spare.empty()) {
      L = spare.front();
      spare.erase(spare.begin());
    } else {
      L = new MgLevel;
    }
    L->reset();
    lv.push_back(L);
    return L;
  }
  void drop_spare() {
    for (MgLevel* l : spare) delete l;
    spare.clear();
  }
  ~Multigrid() {
    join_prepared();
    clear();
    drop_spare();
    if (pinned) (void)hipHostFree(pinned);
    if (pinned_pat) (void)hipHostFree(pinned_pat);
    if (pat_ev) (void)hipEventDestroy(pat_ev);
    for (hipEvent_t& e : pre_piece)
      if (e) (void)hipEventDestroy(e);
    if (map_ev) (void)hipEventDestroy(map_ev);
    if (pre_queued) (void)hipEventDestroy(pre_queued);
    if (pre_done) (void)hipEventDestroy(pre_done);
    if (bound_ev) (void)hipEventDestroy(bound_ev);
    if (lam_ev) (void)hipEventDestroy(lam_ev);
    for (hipEvent_t e : lvl_ev) (void)hipEventDestroy(e);
  }
  void clear() {
    spare.insert(spare.begin(), lv.begin(), lv.end());
    lv.clear();
    tail_from = -1;
    tail_packed = -1;
    m_bound = 0.0;
    ready = false;
  }
};

struct Context {
  int device = 0;
  hipStream_t stream = nullptr;
  hipStream_t aux = nullptr;  // copies that may overlap the main stream (created on first use)
  int cu_count = 0;

  // mesh (host copies kept for boundary loads)
  int nn = 0;
  int64_t ne = 0;
  std::vector<double> h_xy;
  std::vector<int> h_tri;
  DBuf<double> xy, E_elem, nu_elem;
  DBuf<int> tri;
  bool has_E_elem = false;
  bool has_nu_elem = false;  // per-element Poisson ratio (only with a per-element modulus)
  double E = 1.0, nu = 0.3, thick = 1.0;
  int plane_strain = 0;
  Mat3 D{};  // uniform constitutive matrix (used when there is no per-element modulus)

  DBuf<int> ne_ptr, ne_list;

  // pattern (block CSR) and slot map
  bool have_pattern = false;
  bool mesh_pattern = false;  // pattern came from the mesh (slot map valid)
  int nb = 0;                 // block rows
  int n = 0;                  // scalar rows (== 2*nb for mesh matrices)
  int64_t nblk = 0;           // real blocks
  DBuf<int> brow_ptr, bcol, bdiag;
  DBuf<int> slot_ptr, slot_src;

  // SELL storage
  std::vector<int> h_segments;   // block-row boundaries slices may not cross (batches)
  std::vector<int> h_seg_slice;  // first slice of each segment (+ end)
  DBuf<int> sell_row0, sell_nrow;
  int nslices = 0;
  int64_t sell_blocks = 0;
  std::vector<int64_t> h_sell_off;
  DBuf<int64_t> sell_off;
  DBuf<int> sell_col;
  DBuf<double> sell_val;   // 4 per block, planar per (slice, k): see sell_vidx
  DBuf<int64_t> sell_pos;  // CSR block -> SELL slot
  DBuf<float> sell_val32;  // FP32 copy (v00 v01 v10 v11 at 4*pos) for the mixed-precision path
  bool vals32_valid = false;

  // generic CSR upload: scalar entry -> SELL value index (host side, for download)
  std::vector<int64_t> h_generic_map;
  bool generic = false;

  // vectors
  DBuf<double> b, x, r, p, Ap, inv_diag, work;
  DBuf<float> r32, p32, Ap32, x32, invd32;
  DBuf<double> z64, ax64;  // the FP64 flexible CG's z and A x (kept across solves)
  bool setup_done = false;

  // batch of independent plates (batched.hip)
  int batch_count = 0;
  DBuf<int> batch_seg_slice, batch_order, batch_queue;
  std::vector<int> h_batch_order;  // plates by decreasing estimated cost
  // teams of the batched solve: member map, single-plate list, arrival counters, error flag,
  // and the teams' per-slice partials (3 x slices)
  DBuf<int4> batch_team;
  DBuf<int> batch_single, batch_team_cnt, batch_team_err;
  DBuf<double> batch_gpart;
  DBuf<cstfem_batch_result> batch_results;
  // the batched solve's captured graph, replayed while its launch is unchanged (key: its
  // precision, shape and kernel arguments)
  hipGraph_t batch_graph = nullptr;
  hipGraphExec_t batch_exec = nullptr;
  std::vector<unsigned char> batch_graph_key;
  void drop_batch_graph() {
    if (batch_exec) (void)hipGraphExecDestroy(batch_exec);
    if (batch_graph) (void)hipGraphDestroy(batch_graph);
    batch_exec = nullptr;
    batch_graph = nullptr;
    batch_graph_key.clear();
  }

  // constrained DOFs (union over apply_dirichlet calls since the last assembly): one byte
  // per DOF, allocated with the vectors; fixed_valid is false until the first elimination
  // after an assembly or upload has cleared it
  DBuf<unsigned char> fixed;
  bool fixed_valid = false;
  bool pattern_symmetric = true;  // structurally symmetric block pattern (mesh patterns are)
  int max_row_blocks = 0;         // blocks of the longest block row
  // Dirichlet elimination workspace, grown on demand and kept: sorted constrained DOFs and
  // values, the DOF and block-row bitmaps (zero between calls) and the affected-row list
  DBuf<int> dir_dofs, dir_rows, dir_flag;  // dir_flag: a word per block row (zero between calls)
  DBuf<double> dir_vals, dir_uval;  // sorted values; prescribed value by DOF (n entries)
  DBuf<unsigned> dir_bits;
  // device address of the allocation-free elimination's scratch array on this GPU
  unsigned char* na_scratch = nullptr;
  Mpnn mpnn;
  Multigrid mg;

  // grow-only integer scratch of the pattern builders and scans, one buffer per call site
  // (slots in assembly.hip), kept so that rebuilding a pattern allocates and frees nothing
  DBuf<int> tmp_i[12];

  // staging of cstfem_element_stiffness: 6 coordinates and 36 entries per triangle
  DBuf<double> ke_io;

  // traction workspace (loads.hip): packed edge/node index lists, edge geometry and the
  // tractions at the Gauss points, grown on demand and kept
  DBuf<int> trac_i;
  DBuf<double> trac_d;
  DBuf<double> body_w;  // element weights t*|A|/3 of the body force
  // constant-traction groups already applied on this mesh: their packed index lists on the
  // device, their edges checked (each on exactly one triangle).  A repeated group skips the
  // host packing, the upload and the wait for the check.  Dropped when the mesh changes.
  struct TractionGroup {
    std::vector<int> edges;
    DBuf<int> idx;
    int U = 0;
  };
  std::vector<std::unique_ptr<TractionGroup>> trac_groups;
  static constexpr size_t kTractionGroups = 16;

  // reduction scratch
  DBuf<double> partials;   // 4 * max_blocks
  DBuf<PcgScalars> scal;
  PcgScalars* h_scal = nullptr;  // pinned
  DBuf<int> err_flag;            // [0] = min bad index (INT_MAX = none)
  int red_blocks = 0;            // grid used by reduction kernels
  // page-locked staging of the small host<->device transfers of loads and Dirichlet (lists,
  // counters): copies from pageable memory cost the host a staging copy and a wait each.
  // Grown on demand; every user synchronises the stream before it returns, so the next
  // user finds no copy in flight.
  unsigned char* h_stage = nullptr;
  size_t h_stage_n = 0;
  unsigned char* stage(size_t bytes);

  ~Context();
};

// thread-local error message of the C ABI
void set_error(const std::string& m);

// helpers implemented in the .hip files
// exclusive prefix sums on the library stream; with total, waits for them and returns the sum
void exclusive_scan_int(Context& c, const int* in, int* out, int64_t count, int64_t* total);
// the context's integer scratch buffer `slot` (0..11), grown to at least count elements
int* tmp_ints(Context& c, int slot, size_t count);
int grid_for(int64_t threads, int block, int max_grid);

// kernels-level entry points (assembly.hip, pcg.hip, ...)
void build_node_elements(Context& c);
void build_pattern(Context& c);
void build_sell(Context& c);
void build_slot_map(Context& c);
void assemble_values(Context& c);
void element_stiffness(Context& c, int64_t count, const double* xy, const double* D, double t,
                       double* Ke);
void download_values(Context& c, double* vals);
void apply_dirichlet_device(Context& c, int m, const int* dofs, const double* values);
// the same elimination with the fixed per-GPU scratch array instead of context buffers
void apply_dirichlet_noalloc(Context& c, int m, const int* dofs, const double* values);
// looks up (once per device) the scratch array's device address for this context's GPU
void noalloc_attach(Context& c);
void add_body_force_device(Context& c, double bx, double by);
// where the traction at a Gauss point comes from: a constant (p0, p1), the Kirsch field with
// (sigma_inf, a) = (p0, p1), or a caller's function
struct TractionSource {
  enum Kind { kConstant, kKirsch, kFunction } kind = kConstant;
  double p0 = 0.0, p1 = 0.0;
  cstfem_traction_fn fn = nullptr;
  void* user = nullptr;
};
void add_traction_device(Context& c, int nedge, const int* edges, const TractionSource& src);
void kirsch_traction(double sigma_inf, double a, double x, double y, double nx, double ny,
                     double t[2]);
void setup_solver(Context& c);
void spmv_fp64(Context& c, const double* x, double* y);
// a'b over the 2*nb entries of two device vectors (fixed-order reduction), synchronous
double dot_fp64(Context& c, const double* a, const double* b);
int solve_cg_fp64(Context& c, const cstfem_solver_options& o, cstfem_solver_stats& st,
                  double* history);
int solve_refined(Context& c, const cstfem_solver_options& o, cstfem_solver_stats& st, bool fp32,
                  bool warm = false);
int solve_fcg_learned(Context& c, const cstfem_solver_options& o, cstfem_solver_stats& st,
                      bool fp32);
void ensure_vectors(Context& c);
void ensure_vals32(Context& c);
// multilevel preconditioner (mg.hip)
void mg_setup(Context& c, int max_levels, int coarse_max, int fine_smoother, int kdepth);
// starts the level-0 aggregation of the current mesh pattern (node-graph download on the
// auxiliary stream, greedy pass on a host thread); the next mg_setup takes its result
void mg_prepare(Context& c);
// starts the level-0 aggregation of the current mesh pattern now (mg_prepare's work; the
// pattern build calls it right after the node graph once mg_prepare has been called)
void mg_prepare_start(Context& c);
// waits for a started level-0 aggregation and drops it (before the pattern changes)
void mg_drop_prepared(Context& c);
// z = B r on the library stream (level-0 vectors of 2*nb doubles); kernels return at once
// when sc is given and its status is no longer running
// With zq_part (and the learned smoother), the last kernel also leaves the block partials of
// z'zq there and the count is returned; 0 when they were not formed
int mg_enqueue_apply(Context& c, const double* r, double* z, const PcgScalars* sc,
                     const double* zq = nullptr, double* zq_part = nullptr);
// z = P C(P^T r) with C the hierarchy's coarse correction (no level-0 smoothing)
void mg_enqueue_coarse(Context& c, const double* r, double* z);
int64_t mg_level_info(Context& c, int level, int* n, int* nagg, double* lam);
int mg_levels(Context& c);
void mg_tail_info(Context& c, int* first_level, int* packed);
// switches the level-0 smoother (0 learned M, 1 Chebyshev; its spectral bound is computed on
// first use)
void mg_set_fine_smoother(Context& c, int smoother);
void mg_download_agg(Context& c, int level, int* out, int count);
void mg_download_matrix(Context& c, int level, int* row_ptr, int* cols, double* vals);
int mpnn_num_params();
void mpnn_init(Context& c, int seed);
void mpnn_set_params(Context& c, const float* theta, int n);
void mpnn_get_params(Context& c, float* theta, int n);
void mpnn_forward(Context& c);
void mpnn_set_pairs(Context& c, int K, const float* R, const float* E);
int mpnn_harvest(Context& c, int k_first, int k_last, int stride);
void mpnn_project_pairs(Context& c);
void mpnn_add_white_pairs(Context& c, int count, int seed);
double mpnn_loss_grad(Context& c);
double mpnn_train_step(Context& c, float lr);
void mpnn_download(Context& c, int what, float* out, int64_t count);
double profile_kernel(Context& c, int kind, int reps, int64_t bytes);
void element_stresses(Context& c, double* sig, double* vm, double* szz, double* cxy, double* area);
void kirsch_metrics(Context& c, double sigma_inf, double a, cstfem_kirsch_result* out);
void batch_define(Context& c, int count, const int* node_off);
// iterates a batched solve records: x after iteration first + j at x + j * stride
struct BatchSnapshots {
  double* x;
  int first, count;
  int64_t stride;
};
double batch_solve(Context& c, const cstfem_solver_options& o, const cstfem_batch_config& cfg,
                   cstfem_batch_result* results, const BatchSnapshots* snaps = nullptr);
int batch_harvest(Context& c, const cstfem_solver_options& o, int k_first, int k_last);
cstfem_batch_config batch_autotune(Context& c, const cstfem_solver_options& o,
                                   const char* cache_path, double* best_seconds);

}  // namespace cstfem_hip
