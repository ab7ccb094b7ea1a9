/*
 * cstfem_hip.h -- C ABI of the MI355X (gfx950) assemble-and-solve path for 2D linear-elastic
 * constant-strain triangles.
 *
 * Conventions (kept from the reference, include/cstfem headers):
 *   - indices are 32-bit int, DOF = 2*node + component (assembly.hpp:10);
 *   - the stiffness pattern is the reference's CSR (csr.hpp:14-32): row_ptr[n+1], sorted
 *     cols[nnz], rows 2n and 2n+1 share one column list [2m, 2m+1 for m in adj(n)];
 *   - the caller owns every host pointer; the context owns all device memory;
 *   - every call returns a status: 0 ok, 2 not converged, negative = error class below;
 *     cstfem_last_error() gives the message (thread-local), with the reference's message
 *     prefixes ("cg_jacobi: non-positive diagonal at row", "apply_dirichlet: ...").
 * One context per host thread; contexts share nothing.
 */
#ifndef CSTFEM_HIP_H
#define CSTFEM_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  CSTFEM_OK = 0,
  CSTFEM_NOT_CONVERGED = 2,
  CSTFEM_ERR_ARGUMENT = -1,
  CSTFEM_ERR_DEVICE = -2,
  CSTFEM_ERR_BAD_DIAGONAL = -3,   /* std::runtime_error in the reference */
  CSTFEM_ERR_MISSING_ENTRY = -4,  /* std::logic_error in the reference */
  CSTFEM_ERR_OUT_OF_RANGE = -5,   /* std::runtime_error */
  CSTFEM_ERR_ALLOC = -6,
  CSTFEM_ERR_TRACTION_EDGE = -7,  /* std::runtime_error */
  CSTFEM_ERR_CONFLICT = -8,       /* std::runtime_error: conflicting Dirichlet values */
  CSTFEM_ERR_ORDER = -9           /* std::logic_error: e.g. solve() before setup() */
};

enum { CSTFEM_PLANE_STRESS = 0, CSTFEM_PLANE_STRAIN = 1 };

/* Solver precision / method of cstfem_solve(). */
enum {
  CSTFEM_METHOD_CG_FP64 = 0,        /* Jacobi-PCG in FP64, the reference recurrence; stops on
                                       the recursive residual (pcg_jacobi, solver.cpp:109) */
  CSTFEM_METHOD_IR_FP32 = 1,        /* FP32 Jacobi-PCG (FP32 matrix and vectors, FP64 dots,
                                       FP64 solution) refined by FP64 true residuals; stops when
                                       the FP64 true residual is below rtol */
  CSTFEM_METHOD_FCG_LEARNED = 2,    /* flexible CG (one kept direction) with the learned
                                       preconditioner: FP64 matrix and vectors, M in FP32;
                                       falls back to Jacobi-CG on breakdown or stall; stops on
                                       the FP64 true residual */
  CSTFEM_METHOD_CG_FP64_REFINED = 3,/* the FP64 recurrence plus FP64 true-residual replacement
                                       until the true residual is below rtol */
  CSTFEM_METHOD_FCG_LEARNED_FP32 = 4 /* the flexible CG with FP32 matrix and vectors (FP64 dots
                                       and solution), FP64 true-residual refreshes that keep the
                                       search direction */
};

typedef struct cstfem_ctx cstfem_ctx;

/* Mirrors SolverOptions (solver.hpp:93-97) plus device-path knobs. */
typedef struct {
  double rtol;            /* default 1e-10 */
  int max_iter;           /* default 20000 */
  int record_history;     /* residual_history of ||r_k||/||b|| (k = 0..iterations) */
  int instrument;         /* fill the breakdown fields (kernel-class times) */
  int method;             /* CSTFEM_METHOD_* */
  double inner_rtol;      /* refined methods: refresh r from the FP64 true residual when the
                             recursive residual drops by this factor since the last refresh
                             (0 = default: 1e-3 in FP32; in FP64 only when the recursive
                             residual reaches rtol, where the reference stops) */
  int max_outer;          /* refined methods: max refreshes (0 = default 200) */
  int refresh;            /* refined methods, after r := b - A x:  0 restart (p := D^-1 r, the
                             default), 1 keep the search direction after a corrective line
                             search that restores r'p = 0, 2 keep the search direction */
} cstfem_solver_options;

/* Mirrors SolverStats (solver.hpp:99-107) and SolveBreakdown (solver.hpp:17-35). */
typedef struct {
  int iterations;          /* total inner iterations (FP64 CG: the reference's count) */
  int converged;
  int breakdown;           /* 1 if p'Ap <= 0 stopped CG */
  int zero_rhs;
  int outer_iterations;    /* IR / FCG outer steps; 0 for plain CG */
  int fallback;            /* FCG: 0 none; on stagnation or breakdown 1 = fell back to
                              Jacobi-CG, 2 = to the multilevel cycle with the Chebyshev fine
                              smoother (learned multilevel only) */
  int warning;             /* 1 if converged but true residual > 10*rtol */
  int pad_;
  double rel_residual;     /* recursive ||r||/||b|| at exit */
  double true_rel_residual;/* ||b - A x|| / ||b|| in FP64 */
  double breakdown_pap;
  /* breakdown (seconds; the reference's field names) */
  double setup, spmv_local, dot_local, vector_ops, total;
  long long spmv_calls, reductions;
} cstfem_solver_stats;

/* ---- context ------------------------------------------------------------------------ */
const char* cstfem_last_error(void);
int cstfem_ctx_create(int device, cstfem_ctx** out);
int cstfem_ctx_destroy(cstfem_ctx* ctx);
/* Device properties for reports: name (NUL-terminated, <= len), CU count. */
int cstfem_device_info(cstfem_ctx* ctx, char* name, int len, int* cu_count);

/* ---- mesh and material (replaces the Mesh argument of assembly.hpp:15,22) -------------- */
/* xy[2*nn], tri[3*ne] counter-clockwise, E_elem[ne] or NULL (then E is uniform, as the
 * reference's constitutive_matrix(E, nu, plane), element.cpp:7-15). */
int cstfem_upload_mesh(cstfem_ctx* ctx, int nn, const double* xy, int64_t ne, const int* tri);
int cstfem_set_material(cstfem_ctx* ctx, double E, double nu, int plane, double thickness,
                        const double* E_elem);
/* The reference's assemble_stiffness(mesh, D, thickness, K) form: a uniform row-major 3x3
 * constitutive matrix D[9] (assembly.hpp:22); clears any per-element modulus. */
int cstfem_set_constitutive(cstfem_ctx* ctx, const double* D, double thickness);

/* ---- element stiffness ---------------------------------------------------------------- */
/* Replaces cst_stiffness(xy, D, thickness) (element.hpp:27, element.cpp:37-55) for `count`
 * triangles at once: xy[6*count] holds (x1, y1, x2, y2, x3, y3) per triangle (TriCoords),
 * D[9] the row-major constitutive matrix (constitutive_matrix, element.cpp:7-15), and
 * Ke[36*count] receives each row-major 6x6 Ke = t |A| B^T D B, upper triangle computed as
 * the reference computes it and mirrored, so equal to it bit for bit. */
int cstfem_element_stiffness(cstfem_ctx* ctx, int64_t count, const double* xy, const double* D,
                             double thickness, double* Ke);

/* ---- pattern and assembly ------------------------------------------------------------ */
/* Replaces build_stiffness_pattern (assembly.cpp:9-42): builds the pattern on the device,
 * plus the element-to-slot map used by assembly.  *nnz_out = scalar nnz. */
int cstfem_build_pattern(cstfem_ctx* ctx, int64_t* nnz_out);
/* Copies the pattern to the host: row_ptr[2*nn+1], cols[nnz] (either may be NULL). */
int cstfem_download_pattern(cstfem_ctx* ctx, int* row_ptr, int* cols);
/* Replaces assemble_stiffness (assembly.cpp:54-74): zeroes and sums every element's Ke
 * through the slot map, in element order per entry (no atomics). */
int cstfem_assemble(cstfem_ctx* ctx);
/* Copies the values in the reference's CSR order (vals[nnz]). */
int cstfem_download_values(cstfem_ctx* ctx, double* vals);
/* Generic CSR input (the CsrMatrix handed to CgJacobiSolver::setup, solver.cpp:133):
 * n rows, sorted columns.  Replaces any device matrix.  Values are copied. */
int cstfem_upload_csr(cstfem_ctx* ctx, int n, const int* row_ptr, const int* cols,
                      const double* vals);

/* ---- right-hand side and Dirichlet elimination ---------------------------------------- */
/* The global load vector f (driver.cpp:135-148 builds it as zeros plus the loads).  n must
 * equal the system size; b == NULL sets f = 0 on the device without a host copy. */
int cstfem_set_rhs(cstfem_ctx* ctx, const double* b, int n);
int cstfem_download_rhs(cstfem_ctx* ctx, double* b, int n);
/* Adds the consistent nodal load of a uniform body force (bx, by) per unit volume,
 * t*|A|/3 per node of every element, in element order (no reference counterpart). */
int cstfem_add_body_force(cstfem_ctx* ctx, double bx, double by);
/* Replaces add_traction(mesh, group, thickness, tx, ty, f) (bc.cpp:190-253, bc.hpp:73) for
 * one line group given as nedge node pairs edges[2*nedge]: 2-point Gauss per edge, outward
 * normal from the adjacent triangle, accumulated into the device RHS in the reference's edge
 * order.  mode 0: constant (p0, p1) = (tx, ty), entirely on the device; mode 1: the Kirsch
 * traction with (p0, p1) = (sigma_inf, a) (kirsch.cpp:40-45), evaluated on the host at the
 * Gauss points the device computed.  An edge not adjacent to exactly one triangle leaves the
 * RHS unchanged and returns CSTFEM_ERR_TRACTION_EDGE. */
int cstfem_add_traction(cstfem_ctx* ctx, int nedge, const int* edges, int mode, double p0,
                        double p1);
/* The reference's TractionFn (bc.hpp:65): traction (tx, ty) at point (x, y) with outward unit
 * normal (nx, ny), called once per add_traction with all 2*nedge Gauss points:
 * pts[4*count] = (x, y, nx, ny) per point, t[2*count] = (tx, ty) to be written. */
typedef void (*cstfem_traction_fn)(void* user, int count, const double* pts, double* t);
/* Replaces add_traction(mesh, group, thickness, const TractionFn&, f) (bc.hpp:71-72): the
 * geometry, quadrature and accumulation run on the device, fn on the host. */
int cstfem_add_traction_fn(cstfem_ctx* ctx, int nedge, const int* edges, cstfem_traction_fn fn,
                           void* user);
/* Replaces apply_dirichlet (bc.cpp:137-166) on the device matrix and RHS.  dofs/values:
 * m prescribed DOFs; a repeated DOF must repeat its value (DirichletDofs::add, bc.cpp:107). */
int cstfem_apply_dirichlet(cstfem_ctx* ctx, int m, const int* dofs, const double* values);
/* The same elimination without any device allocation.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Arguments, results and errors are those of
 * cstfem_apply_dirichlet.  Its workspace is one fixed device array per GPU (13 MiB, held
 * until the process ends) instead of context buffers, which limits a call to systems of at
 * most 2^25 DOFs, at most 2^18 distinct constrained DOFs and at most 2^20 block rows visited
 * (the constrained nodes and their neighbours); a larger call returns CSTFEM_ERR_ARGUMENT and
 * leaves K and b unchanged.  Calls on one GPU share the array and run one at a time. */
int cstfem_apply_dirichlet_noalloc(cstfem_ctx* ctx, int m, const int* dofs,
                                   const double* values);

/* ---- solve (replaces CgJacobiSolver::setup + solve / pcg_jacobi, solver.cpp:109-157) --- */
int cstfem_setup(cstfem_ctx* ctx);
/* x_out[n] (may be NULL: the solution stays on the device); history: NULL or room for
 * max_iter + 1 doubles (plain CG only). */
int cstfem_solve(cstfem_ctx* ctx, const cstfem_solver_options* opt, cstfem_solver_stats* st,
                 double* x_out, double* history);
int cstfem_download_solution(cstfem_ctx* ctx, double* x, int n);
/* y = A x with the device matrix (FP64, per-row summation in column order); replaces
 * CsrMatrix::multiply (csr.cpp:19-26) and CsrOperator::apply (solver.cpp:15-22). */
int cstfem_spmv(cstfem_ctx* ctx, const double* x, double* y, int n);
/* *out = a'b over n entries (n = the system size), summed on the device in a fixed order;
 * replaces SerialReduction::dot (solver.cpp:24-31). */
int cstfem_dot(cstfem_ctx* ctx, const double* a, const double* b, int n, double* out);

/* ---- stresses and Kirsch metrics -------------------------------------------------------- */
/* Replaces compute_element_stresses (postprocess.cpp:35-60) for the last solution on the
 * device: sig[3*ne] = (sxx, syy, txy) per element (element.cpp:57-69, the reference's
 * operation order), von_mises[ne] and szz[ne] (element.cpp:72-81), cxy[2*ne] centroids and
 * area[ne]; every output but sig may be NULL.  A per-element modulus is honoured. */
int cstfem_element_stresses(cstfem_ctx* ctx, double* sig, double* von_mises, double* szz,
                            double* cxy, double* area);
/* Mirrors KirschMetrics (postprocess.hpp) */
typedef struct {
  double l2_rel_stress_error;       /* centroid stresses vs the analytic field, area-weighted */
  double l2_rel_displacement_error; /* degree-2 rule */
  double energy_rel_error;          /* complementary energy norm, degree-2 rule */
  int scf_element;                  /* element whose centroid is nearest to (0, a) */
  int pad_;
  double scf_x, scf_y, scf_numerical, scf_analytical_centroid, scf_theoretical;
} cstfem_kirsch_result;
/* Replaces kirsch_metrics (postprocess.cpp:81-150) for the last solution: uniform modulus
 * (cstfem_set_material without E_elem), sigma_inf and hole radius a. */
int cstfem_kirsch_metrics(cstfem_ctx* ctx, double sigma_inf, double a, cstfem_kirsch_result* out);

/* ---- batches of independent plates (no reference counterpart) ------------------------- */
/* Per-element modulus and Poisson ratio (plates of a batch differ in both); D_e is
 * constitutive_matrix(E_e, nu_e, plane) (element.cpp:7-15). */
int cstfem_set_material_fields(cstfem_ctx* ctx, const double* E_elem, const double* nu_elem,
                               int plane, double thickness);
/* The uploaded mesh is the concatenation of `count` plates, plate k owning nodes
 * [node_off[k], node_off[k+1]); no element may join two plates.  Call after
 * cstfem_build_pattern and before cstfem_assemble (the matrix layout is re-aligned to the
 * plates). */
int cstfem_batch_define(cstfem_ctx* ctx, int count, const int* node_off);
/* Launch shape of the persistent batched solve: wavefronts per workgroup (4, 8, 16),
 * workgroups per CU (1..4, at most what fits), and the most workgroups one plate may take
 * (max_team: 1, 2, 4, 8 or 16; 0 = 16).  With max_team > 1 the largest plates are each
 * solved by a team of co-resident workgroups that split its slices and meet at barriers, so
 * that one big plate no longer sets the makespan.  Results do not depend on the shape: the
 * reductions sum per-slice partials in one fixed order whoever computed them. */
typedef struct {
  int waves_per_group;
  int groups_per_cu;
  int max_team;
} cstfem_batch_config;
typedef struct {
  int iterations;
  int converged;            /* FP64 true residual below rtol */
  int refinements;          /* FP64 true-residual refreshes */
  int status;               /* 1 converged, 2 breakdown, 3 max_iter / no progress */
  double rel_residual;      /* recursive ||r||/||b|| at exit */
  double true_rel_residual; /* FP64 ||b - A x|| / ||b|| at the last refresh */
} cstfem_batch_result;
/* Solves every plate in one persistent launch (method CSTFEM_METHOD_CG_FP64[_REFINED] or
 * CSTFEM_METHOD_IR_FP32; every plate is verified by its FP64 true residual).  cfg NULL = the
 * default shape (8 waves, 1 group per CU, teams of up to 16).  results[count]; x_out[n] may be NULL; *seconds =
 * device time of the launch.  Returns 2 if any plate did not converge. */
int cstfem_batch_solve(cstfem_ctx* ctx, const cstfem_solver_options* opt,
                       const cstfem_batch_config* cfg, cstfem_batch_result* results,
                       double* x_out, double* seconds);
/* Times every launch shape on the current batch, checks that all give identical results,
 * returns the fastest; cached per device/batch shape/method in cache_path (NULL = no cache). */
int cstfem_batch_autotune(cstfem_ctx* ctx, const cstfem_solver_options* opt,
                          const char* cache_path, cstfem_batch_config* best,
                          double* best_seconds);

/* ---- learned preconditioner (no reference counterpart) ------------------------------- */
/* A message-passing network over the mesh graph predicts the 2x2 blocks of a symmetric
 * sparse approximate inverse on the stiffness pattern,
 * M_ij = S_i [[a, b], [b, d]] S_j + c D_i^-1 K_ij D_j^-1 (S = D^-1/2, D = diag(K));
 * CSTFEM_METHOD_FCG_LEARNED runs flexible CG with M and falls back to Jacobi-CG from the
 * current iterate when it stagnates or breaks down.  Parameters are FP32 (a flat vector of
 * cstfem_mpnn_num_params() floats); the network computes in BF16 with FP32 accumulation.
 * All calls below need an assembled, constrained mesh system and cstfem_setup. */
int cstfem_mpnn_num_params(void);
/* Random-normal hidden layers (He scaling) from `seed`, decoder = Jacobi (M = D^-1). */
int cstfem_mpnn_init(cstfem_ctx* ctx, int seed);
int cstfem_mpnn_set_params(cstfem_ctx* ctx, const float* theta, int n);
int cstfem_mpnn_get_params(cstfem_ctx* ctx, float* theta, int n);
/* Node features, network forward and M for the current matrix (before a FCG solve). */
int cstfem_mpnn_build(cstfem_ctx* ctx);
/* Training pairs from the current system: x* by FP64 CG to 1e-12, then e = x* - x_k and
 * r = K e after plain FP64 Jacobi-CG iterations k = k_first, k_first + stride, ... <= k_last.
 * *npairs = number of pairs kept on the device. */
int cstfem_mpnn_harvest(cstfem_ctx* ctx, int k_first, int k_last, int stride, int* npairs);
/* Training pairs from a batch of plates (cstfem_batch_define): x* of every plate by the
 * batched solve with opt (an FP64 method), then every plate's plain Jacobi-CG from x = 0 (the
 * reference recurrence) for k_last iterations, recording e_k = x* - x_k and r_k = K e_k after
 * iterations k = k_first..k_last; pair row k - k_first holds iteration k of every plate.
 * Leaves x = x_{k_last}.  *npairs = k_last - k_first + 1. */
int cstfem_batch_harvest(cstfem_ctx* ctx, const cstfem_solver_options* opt, int k_first,
                         int k_last, int* npairs);
/* Appends `count` white-noise training pairs drawn on the device: e ~ N(0, I) from `seed`,
 * r = K e (FP64 SpMV); existing pairs are kept. */
int cstfem_mpnn_add_white_pairs(cstfem_ctx* ctx, int count, int seed);
/* Replaces every training pair (r, e) by (K e', e') with e' = e - P C(P^T K e): the part of
 * the error that the multilevel hierarchy's coarse correction C (cstfem_mg_setup, K-cycles
 * included, no level-0 smoothing) leaves, which is what the fine-level smoother must reduce. */
int cstfem_mpnn_project_pairs(cstfem_ctx* ctx);
/* Caller-supplied pairs: R, E are K rows of n floats (either may be NULL = zeros). */
int cstfem_mpnn_set_pairs(cstfem_ctx* ctx, int K, const float* R, const float* E);
/* Forward, loss (1/K) sum_k |M r_k - e_k|_K^2 / |e_k|_K^2 and its gradient w.r.t. the
 * parameters (kept on the device, see cstfem_mpnn_download). */
int cstfem_mpnn_loss_grad(cstfem_ctx* ctx, double* loss);
/* cstfem_mpnn_loss_grad followed by one Adam step (beta 0.9 / 0.999, eps 1e-8). */
int cstfem_mpnn_train_step(cstfem_ctx* ctx, float lr, double* loss);
/* what: 0 node features (nb x 16), 1 final hidden state (nb x 64), 2 decoder outputs
 * [a, b, d, c] per stored block (nblk x 4), 3 gradient (num_params), 4 / 5 the training
 * pairs' residuals / errors (npairs x n).  count must match. */
int cstfem_mpnn_download(cstfem_ctx* ctx, int what, float* out, int64_t count);

/* ---- multilevel preconditioner (no reference counterpart) -------------------------------- */
/* Aggregation hierarchy for CSTFEM_METHOD_FCG_LEARNED: greedy aggregation of the node graph,
 * unsmoothed prolongators from the rigid-body modes (3 coarse DOFs per aggregate), Galerkin
 * coarse matrices on the device, a dense coarsest solve; coarse levels 1..kcycle_levels are
 * solved by a K-cycle (two flexible CG steps each), deeper ones by one V-cycle (3 is the
 * fastest on the 4M-DOF plate: the same iterations as 16 with fewer coarse launches).
 * fine_smoother: 0 the learned M (cstfem_mpnn_build first), 1 Chebyshev (degree 2) on
 * D^-1 K.  Levels are added until a level has <= coarse_max DOFs or max_levels is reached
 * (the coarsest must have <= 4096 DOFs).  Call after cstfem_setup; a new setup drops the
 * hierarchy.  While a hierarchy exists, the FP64 flexible CG uses it.  With
 * fine_smoother = 0 and M built, the set-up estimates lambda_max(M K) (12 power steps); above
 * 1.5 (or not finite) the learned smoother is replaced by Chebyshev for this hierarchy. */
int cstfem_mg_setup(cstfem_ctx* ctx, int max_levels, int coarse_max, int fine_smoother,
                    int kcycle_levels);
/* Starts the set-up's first step ahead of time: the greedy aggregation of the fine node graph,
 * which depends on the pattern only.  The graph is copied to the host on a second stream and
 * aggregated on a worker thread, while the caller goes on with assembly, loads, Dirichlet,
 * cstfem_setup and cstfem_mpnn_build; the next cstfem_mg_setup waits for it and uses it.
 * Call after cstfem_build_stiffness_pattern; a new pattern (or mesh, or CSR upload) drops a
 * started aggregation.  From the first call on, every later cstfem_build_stiffness_pattern of
 * the context starts it by itself as soon as the node graph exists (before the pattern's SELL
 * layout and slot map), and a call for a pattern whose aggregation is under way does nothing.
 * The result is the same as without it. */
int cstfem_mg_prepare(cstfem_ctx* ctx);
int cstfem_mg_levels(cstfem_ctx* ctx, int* levels);
/* DOFs, aggregates (0 on the coarsest), stored entries and the upper end of the smoothing
 * interval (lambda_max estimate of D^-1 K; 0 where no Chebyshev smoother runs) of one level.
 * On level 0 with a learned smoother it is the set-up's estimate of lambda_max(M K). */
int cstfem_mg_level_info(cstfem_ctx* ctx, int level, int* n, int* nagg, int64_t* nnz,
                         double* lam);
/* The fused single-workgroup tail of the cycle: the first level it covers (-1: none) and
 * whether the set-up packed its operands into one array (1: one cycle level over the dense
 * coarsest, every load issued at the kernel's start) or it reads them level by level (0).
 * Both forms give the same bits. */
int cstfem_mg_tail(cstfem_ctx* ctx, int* first_level, int* packed);
/* Aggregate of every node of `level` (count = nodes of the level). */
int cstfem_mg_download_aggregates(cstfem_ctx* ctx, int level, int* agg, int count);
/* Coarse matrix of a level >= 1 as scalar CSR: row_ptr[n+1], cols/vals[nnz]. */
int cstfem_mg_download_matrix(cstfem_ctx* ctx, int level, int* row_ptr, int* cols,
                              double* vals);
/* z = B r: one application of the preconditioner (host vectors of n doubles). */
int cstfem_mg_apply(cstfem_ctx* ctx, const double* r, double* z, int n);

/* ---- measurement (no reference counterpart) ------------------------------------------ */
/* Average seconds of `reps` back-to-back launches of one hot kernel on the library's
 * stream, timed with HIP events; solver state is untouched.  kind: 0 fused SpMV + p'Ap,
 * 1 plain SpMV, 2 CG update (axpys + dots), 3 CG direction, 4 copy of `bytes` bytes,
 * (achievable HBM bandwidth), 5 FP32 SpMV + p'Ap; flexible CG in FP64: 6 SpMV + p'q, p'r,
 * 7 preconditioner z = M r + z'q (after cstfem_mpnn_build), 8 update, 9 direction;
 * 10 read of `bytes` bytes (achievable read-only HBM bandwidth).  Kinds 4 and 10 need no
 * matrix. */
int cstfem_profile_kernel(cstfem_ctx* ctx, int kind, int reps, int64_t bytes,
                          double* avg_seconds);

#ifdef __cplusplus
}
#endif
#endif
