/*
 * cst_oracle.c -- CPU restatement of the reference's assemble-and-solve path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP path is compared
 * against.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it; the product library never links or calls it.
 *
 * Every function restates one reference function (cited file:line, paths relative to
 * the reference repository root) in plain C99 with the same floating-point operation
 * order, so that with -ffp-contract=off it reproduces the reference bit for bit on
 * x86-64.  Two extensions have no reference counterpart and reduce to the reference
 * exactly when unused:
 *   - a per-element Young's modulus (E_elem != NULL): D_e = constitutive(E_e, nu);
 *   - a consistent CST body-force load t*|A|/3*(bx, by) per node of every element.
 *
 * Pinning (what checks each function; see tests/test_oracle.py):
 *   oc_constitutive      closed form of test_element.cpp:121-136 (1e-15)
 *   oc_stiffness         hand KAT test_element.cpp:60-84, rigid modes, symmetry,
 *                        and bitwise equality with the compiled reference (oracle/_ref)
 *   oc_build_pattern     nnz = 4(nodes + 2 edges), sortedness, bitwise vs oracle/_ref
 *   oc_assemble          rigid modes <= 1e-12 and bitwise vs oracle/_ref
 *   oc_spmv              dense product on small matrices, bitwise vs oracle/_ref
 *   oc_apply_dirichlet   3x3 KAT test_bc.cpp:63-78 (exact) and bitwise vs oracle/_ref
 *   oc_add_traction      resultant = t * traction * length (test_bc.cpp:107-126)
 *   oc_add_body_force    resultant = t * area * (bx, by)   (project extension)
 *   oc_pcg_jacobi        dense Cholesky (test_pcg.cpp:75-102), patch tests, Kirsch
 *                        L0/L1 iteration counts 148 / 295 and bitwise vs oracle/_ref
 *   oc_element_stresses  linear-field stress recovery (closed form)
 *   oc_kirsch_*          golden Kirsch metrics of docs/verification_results.json
 *   oc_pcg_jacobi_mt     the same recurrence on several threads (OpenMP), for the timed CPU
 *                        baseline only: its dot products sum per-thread partials in thread
 *                        order, so its rounding differs from oc_pcg_jacobi; the tests check
 *                        that it converges to the same solution (1e-8)
 */
#include <math.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define OC_OK 0
#define OC_NOT_CONVERGED 2
#define OC_ERR_BAD_DIAGONAL (-3)
#define OC_ERR_MISSING_ENTRY (-4)
#define OC_ERR_OUT_OF_RANGE (-5)
#define OC_ERR_ALLOC (-6)
#define OC_ERR_TRACTION_EDGE (-7)

/* element.cpp:7-15 */
void oc_constitutive(double E, double nu, int plane_strain, double D[9]) {
  if (!plane_strain) {
    const double c = E / (1.0 - nu * nu);
    const double v[9] = {c, c * nu, 0.0, c * nu, c, 0.0, 0.0, 0.0, c * (1.0 - nu) / 2.0};
    memcpy(D, v, sizeof v);
    return;
  }
  const double c = E / ((1.0 + nu) * (1.0 - 2.0 * nu));
  const double v[9] = {c * (1.0 - nu), c * nu, 0.0, c * nu, c * (1.0 - nu), 0.0,
                       0.0, 0.0, c * (1.0 - 2.0 * nu) / 2.0};
  memcpy(D, v, sizeof v);
}

/* element.cpp:17-19 */
double oc_signed_area(const double xy[6]) {
  return 0.5 * ((xy[2] - xy[0]) * (xy[5] - xy[1]) - (xy[4] - xy[0]) * (xy[3] - xy[1]));
}

/* element.cpp:21-35 */
void oc_B(const double xy[6], double B[18]) {
  const double x1 = xy[0], y1 = xy[1], x2 = xy[2], y2 = xy[3], x3 = xy[4], y3 = xy[5];
  const double b[3] = {y2 - y3, y3 - y1, y1 - y2};
  const double c[3] = {x3 - x2, x1 - x3, x2 - x1};
  const double inv2A = 1.0 / (2.0 * oc_signed_area(xy));
  memset(B, 0, 18 * sizeof(double));
  for (int i = 0; i < 3; ++i) {
    B[0 * 6 + 2 * i] = b[i] * inv2A;
    B[1 * 6 + 2 * i + 1] = c[i] * inv2A;
    B[2 * 6 + 2 * i] = c[i] * inv2A;
    B[2 * 6 + 2 * i + 1] = b[i] * inv2A;
  }
}

/* element.cpp:37-55: upper triangle computed, then mirrored. */
void oc_stiffness(const double xy[6], const double D[9], double t, double K[36]) {
  double B[18], DB[18];
  oc_B(xy, B);
  memset(DB, 0, sizeof DB);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 6; ++j)
      for (int k = 0; k < 3; ++k) DB[i * 6 + j] += D[i * 3 + k] * B[k * 6 + j];
  const double scale = t * fabs(oc_signed_area(xy));
  for (int i = 0; i < 6; ++i)
    for (int j = i; j < 6; ++j) {
      double s = 0.0;
      for (int k = 0; k < 3; ++k) s += B[k * 6 + i] * DB[k * 6 + j];
      K[i * 6 + j] = K[j * 6 + i] = scale * s;
    }
}

/* element.cpp:57-70: strain B*ue, stress D*strain. */
void oc_strain(const double xy[6], const double ue[6], double e[3]) {
  double B[18];
  oc_B(xy, B);
  e[0] = e[1] = e[2] = 0.0;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 6; ++j) e[i] += B[i * 6 + j] * ue[j];
}

static void elem_coords(const int* tri, const double* xy, int64_t e, double out[6]) {
  /* assembly.cpp:44-52 */
  for (int k = 0; k < 3; ++k) {
    const int n = tri[3 * e + k];
    out[2 * k] = xy[2 * n];
    out[2 * k + 1] = xy[2 * n + 1];
  }
}

static int cmp_int(const void* a, const void* b) {
  const int x = *(const int*)a, y = *(const int*)b;
  return (x > y) - (x < y);
}

/* assembly.cpp:9-42.  Two-call protocol: with cols == NULL only row_ptr is written and
 * nnz returned; with cols != NULL the column indices are written too.  Returns nnz or a
 * negative status. */
int64_t oc_build_pattern(int nn, int64_t ne, const int* tri, int* row_ptr, int* cols) {
  int64_t* start = (int64_t*)calloc((size_t)nn + 1, sizeof(int64_t));
  if (!start) return OC_ERR_ALLOC;
  for (int n = 0; n < nn; ++n) start[n + 1] = 1;
  for (int64_t e = 0; e < ne; ++e)
    for (int a = 0; a < 3; ++a) start[tri[3 * e + a] + 1] += 2;
  for (int n = 0; n < nn; ++n) start[n + 1] += start[n];
  int* adj = (int*)malloc((size_t)start[nn] * sizeof(int));
  int64_t* fill = (int64_t*)malloc((size_t)nn * sizeof(int64_t));
  int* len = (int*)malloc((size_t)nn * sizeof(int));
  if (!adj || !fill || !len) {
    free(start); free(adj); free(fill); free(len);
    return OC_ERR_ALLOC;
  }
  for (int n = 0; n < nn; ++n) {
    fill[n] = start[n];
    adj[fill[n]++] = n; /* self first (assembly.cpp:12) */
  }
  for (int64_t e = 0; e < ne; ++e) {
    const int* t = &tri[3 * e];
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b)
        if (a != b) adj[fill[t[a]]++] = t[b];
  }
  for (int n = 0; n < nn; ++n) {
    int* a = adj + start[n];
    const int cnt = (int)(start[n + 1] - start[n]);
    qsort(a, (size_t)cnt, sizeof(int), cmp_int);
    int u = 0;
    for (int i = 0; i < cnt; ++i)
      if (i == 0 || a[i] != a[u - 1]) a[u++] = a[i];
    len[n] = u;
  }
  int64_t nnz = 0;
  row_ptr[0] = 0;
  for (int n = 0; n < nn; ++n) {
    const int l = 2 * len[n];
    row_ptr[2 * n + 1] = row_ptr[2 * n] + l;
    row_ptr[2 * n + 2] = row_ptr[2 * n + 1] + l;
  }
  nnz = row_ptr[2 * nn];
  if (cols) {
    int64_t k = 0;
    for (int n = 0; n < nn; ++n)
      for (int c = 0; c < 2; ++c)
        for (int i = 0; i < len[n]; ++i) {
          const int m = adj[start[n] + i];
          cols[k++] = 2 * m;
          cols[k++] = 2 * m + 1;
        }
  }
  free(start); free(adj); free(fill); free(len);
  return nnz;
}

/* csr.cpp:7-12 */
static int csr_find(const int* row_ptr, const int* cols, int row, int col) {
  int lo = row_ptr[row], hi = row_ptr[row + 1];
  while (lo < hi) { /* lower_bound */
    const int mid = lo + (hi - lo) / 2;
    if (cols[mid] < col) lo = mid + 1; else hi = mid;
  }
  return (lo < row_ptr[row + 1] && cols[lo] == col) ? lo : -1;
}

/* assembly.cpp:54-74, with the optional per-element modulus (E_elem may be NULL). */
int oc_assemble(int nn, int64_t ne, const int* tri, const double* xy, const double* E_elem,
                double E, double nu, int plane_strain, double t, const int* row_ptr,
                const int* cols, double* vals, int* bad_row, int* bad_col) {
  (void)nn;
  const int64_t nnz = row_ptr[2 * nn];
  for (int64_t k = 0; k < nnz; ++k) vals[k] = 0.0;
  double D[9];
  oc_constitutive(E, nu, plane_strain, D);
  for (int64_t e = 0; e < ne; ++e) {
    double exy[6], Ke[36];
    elem_coords(tri, xy, e, exy);
    if (E_elem) oc_constitutive(E_elem[e], nu, plane_strain, D);
    oc_stiffness(exy, D, t, Ke);
    int dofs[6];
    for (int k = 0; k < 3; ++k) {
      dofs[2 * k] = 2 * tri[3 * e + k];
      dofs[2 * k + 1] = 2 * tri[3 * e + k] + 1;
    }
    for (int i = 0; i < 6; ++i)
      for (int j = 0; j < 6; ++j) {
        const int pos = csr_find(row_ptr, cols, dofs[i], dofs[j]);
        if (pos < 0) {
          if (bad_row) *bad_row = dofs[i];
          if (bad_col) *bad_col = dofs[j];
          return OC_ERR_MISSING_ENTRY;
        }
        vals[pos] += Ke[i * 6 + j];
      }
  }
  return OC_OK;
}

/* csr.cpp:19-26 */
void oc_spmv(int n, const int* row_ptr, const int* cols, const double* vals, const double* x,
             double* y) {
  for (int i = 0; i < n; ++i) {
    double s = 0.0;
    for (int k = row_ptr[i]; k < row_ptr[i + 1]; ++k) s += vals[k] * x[cols[k]];
    y[i] = s;
  }
}

/* bc.cpp:137-166.  dofs/values: m prescribed DOFs (duplicates must already agree, as
 * DirichletDofs::add enforces at bc.cpp:107-116).  On error *bad holds the DOF. */
int oc_apply_dirichlet(int n, const int* row_ptr, const int* cols, double* vals, double* b,
                       int m, const int* dofs, const double* values, int* bad) {
  char* fixed = (char*)calloc((size_t)n + 1, 1);
  double* ubar = (double*)calloc((size_t)n + 1, sizeof(double));
  if (!fixed || !ubar) { free(fixed); free(ubar); return OC_ERR_ALLOC; }
  for (int k = 0; k < m; ++k) {
    if (dofs[k] < 0 || dofs[k] >= n) {
      if (bad) *bad = dofs[k];
      free(fixed); free(ubar);
      return OC_ERR_OUT_OF_RANGE;
    }
    fixed[dofs[k]] = 1;
    ubar[dofs[k]] = values[k];
  }
  for (int i = 0; i < n; ++i) {
    if (fixed[i]) {
      const int kd = csr_find(row_ptr, cols, i, i);
      const double diag = kd >= 0 ? vals[kd] : 0.0;
      if (!(diag > 0.0)) {
        if (bad) *bad = i;
        free(fixed); free(ubar);
        return OC_ERR_BAD_DIAGONAL;
      }
      for (int k = row_ptr[i]; k < row_ptr[i + 1]; ++k)
        if (cols[k] != i) vals[k] = 0.0;
      b[i] = diag * ubar[i];
    } else {
      for (int k = row_ptr[i]; k < row_ptr[i + 1]; ++k)
        if (fixed[cols[k]]) {
          b[i] -= vals[k] * ubar[cols[k]];
          vals[k] = 0.0;
        }
    }
  }
  free(fixed); free(ubar);
  return OC_OK;
}

/* kirsch.cpp:7-38: analytic Kirsch stress (sxx, syy, txy) at (x, y). */
void oc_kirsch_stress(double sigma_inf, double a, double x, double y, double out[3]) {
  const double r = hypot(x, y), theta = atan2(y, x);
  const double q2 = (a * a) / (r * r), q4 = q2 * q2;
  const double c2 = cos(2.0 * theta), s2 = sin(2.0 * theta);
  const double rr = 0.5 * sigma_inf * (1.0 - q2) + 0.5 * sigma_inf * (1.0 - 4.0 * q2 + 3.0 * q4) * c2;
  const double tt = 0.5 * sigma_inf * (1.0 + q2) - 0.5 * sigma_inf * (1.0 + 3.0 * q4) * c2;
  const double rt = -0.5 * sigma_inf * (1.0 + 2.0 * q2 - 3.0 * q4) * s2;
  const double c = cos(theta), n = sin(theta);
  out[0] = rr * c * c + tt * n * n - 2.0 * rt * n * c;
  out[1] = rr * n * n + tt * c * c + 2.0 * rt * n * c;
  out[2] = (rr - tt) * n * c + rt * (c * c - n * n);
}

/* kirsch.cpp:47-60 */
void oc_kirsch_displacement(double sigma_inf, double a, double E, double nu, int plane_strain,
                            double x, double y, double out[2]) {
  const double mu = E / (2.0 * (1.0 + nu));
  const double kappa = !plane_strain ? (3.0 - nu) / (1.0 + nu) : 3.0 - 4.0 * nu;
  const double r = hypot(x, y), th = atan2(y, x);
  const double a2 = a * a, a4 = a2 * a2, r3 = r * r * r;
  const double ur = sigma_inf / (8.0 * mu) *
                    ((kappa - 1.0) * r + 2.0 * a2 / r +
                     2.0 * (r + (kappa + 1.0) * a2 / r - a4 / r3) * cos(2.0 * th));
  const double ut = -sigma_inf / (4.0 * mu) * (r + (kappa - 1.0) * a2 / r + a4 / r3) * sin(2.0 * th);
  const double c = cos(th), s = sin(th);
  out[0] = ur * c - ut * s;
  out[1] = ur * s + ut * c;
}

static uint64_t edge_key(int a, int b) {
  if (a > b) { const int t = a; a = b; b = t; }
  return ((uint64_t)(uint32_t)a << 32) | (uint32_t)b;
}

static int cmp_u64(const void* a, const void* b) {
  const uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
  return (x > y) - (x < y);
}

/* bc.cpp:190-246.  mode 0: uniform traction (p0, p1) = (tx, ty);
 * mode 1: Kirsch traction sigma.n with (p0, p1) = (sigma_inf, a) (kirsch.cpp:40-45).
 * edges: nedge pairs of node indices of the boundary group. */
int oc_add_traction(int nn, int64_t ne, const int* tri, const double* xy, int nedge,
                    const int* edges, double t, int mode, double p0, double p1, double* f,
                    int* bad_edge) {
  (void)nn;
  /* sorted keys of the group's edges with adjacency counts and opposite nodes */
  uint64_t* keys = (uint64_t*)malloc((size_t)(nedge > 0 ? nedge : 1) * sizeof(uint64_t));
  int* count = (int*)calloc((size_t)(nedge > 0 ? nedge : 1), sizeof(int));
  int* opp = (int*)malloc((size_t)(nedge > 0 ? nedge : 1) * sizeof(int));
  if (!keys || !count || !opp) { free(keys); free(count); free(opp); return OC_ERR_ALLOC; }
  for (int k = 0; k < nedge; ++k) keys[k] = edge_key(edges[2 * k], edges[2 * k + 1]);
  qsort(keys, (size_t)nedge, sizeof(uint64_t), cmp_u64);
  int nu = 0;
  for (int k = 0; k < nedge; ++k)
    if (k == 0 || keys[k] != keys[nu - 1]) keys[nu++] = keys[k];
  for (int k = 0; k < nu; ++k) opp[k] = -1;
  for (int64_t e = 0; e < ne; ++e) {
    const int* n = &tri[3 * e];
    for (int k = 0; k < 3; ++k) {
      const uint64_t key = edge_key(n[k], n[(k + 1) % 3]);
      const uint64_t* hit = (const uint64_t*)bsearch(&key, keys, (size_t)nu, sizeof(uint64_t), cmp_u64);
      if (hit) {
        const int idx = (int)(hit - keys);
        opp[idx] = n[(k + 2) % 3];
        ++count[idx];
      }
    }
  }
  const double g = 1.0 / sqrt(3.0);
  const double xi[2] = {-g, g};
  for (int k = 0; k < nedge; ++k) {
    const int a = edges[2 * k], b = edges[2 * k + 1];
    const uint64_t key = edge_key(a, b);
    const int idx = (int)((const uint64_t*)bsearch(&key, keys, (size_t)nu, sizeof(uint64_t), cmp_u64) - keys);
    if (count[idx] != 1) {
      if (bad_edge) *bad_edge = k;
      free(keys); free(count); free(opp);
      return OC_ERR_TRACTION_EDGE;
    }
    const double xa = xy[2 * a], ya = xy[2 * a + 1], xb = xy[2 * b], yb = xy[2 * b + 1];
    const double len = hypot(xb - xa, yb - ya);
    double nx = (yb - ya) / len, ny = -(xb - xa) / len;
    const int c = opp[idx];
    if (nx * (xy[2 * c] - xa) + ny * (xy[2 * c + 1] - ya) > 0.0) { nx = -nx; ny = -ny; }
    for (int q = 0; q < 2; ++q) {
      const double Na = 0.5 * (1.0 - xi[q]), Nb = 0.5 * (1.0 + xi[q]);
      const double x = Na * xa + Nb * xb, y = Na * ya + Nb * yb;
      double tr[2];
      if (mode == 0) {
        tr[0] = p0; tr[1] = p1;
      } else {
        double s[3];
        oc_kirsch_stress(p0, p1, x, y, s);
        tr[0] = s[0] * nx + s[2] * ny;
        tr[1] = s[2] * nx + s[1] * ny;
      }
      const double w = t * 0.5 * len;
      f[2 * a] += w * Na * tr[0];
      f[2 * a + 1] += w * Na * tr[1];
      f[2 * b] += w * Nb * tr[0];
      f[2 * b + 1] += w * Nb * tr[1];
    }
  }
  free(keys); free(count); free(opp);
  return OC_OK;
}

/* Project extension (no reference counterpart): consistent CST load of a uniform body
 * force (bx, by) per unit volume, t*|A|/3 to each node, in element order. */
void oc_add_body_force(int64_t ne, const int* tri, const double* xy, double t, double bx,
                       double by, double* f) {
  for (int64_t e = 0; e < ne; ++e) {
    double exy[6];
    elem_coords(tri, xy, e, exy);
    const double w = t * fabs(oc_signed_area(exy)) / 3.0;
    for (int k = 0; k < 3; ++k) {
      const int n = tri[3 * e + k];
      f[2 * n] += w * bx;
      f[2 * n + 1] += w * by;
    }
  }
}

/* Mirror of SolverStats (solver.hpp:93-107) without the strings. */
typedef struct {
  int iterations;
  int converged;
  int breakdown; /* 1 if p'Ap <= 0 stopped the loop */
  int zero_rhs;
  double rel_residual;
  double true_rel_residual;
  double breakdown_pap;
} oc_stats;

/* solver.cpp:24-31 */
static double dot(int n, const double* a, const double* b) {
  double s = 0.0;
  for (int i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

/* solver.cpp:133-148 (setup) + solver.cpp:35-131 (pcg_iterate + pcg_jacobi).
 * history: NULL or room for max_iter + 1 doubles. Returns OC_OK, OC_NOT_CONVERGED or
 * OC_ERR_BAD_DIAGONAL (row in stats->iterations). */
int oc_pcg_jacobi(int n, const int* row_ptr, const int* cols, const double* vals,
                  const double* b, double* x, double rtol, int max_iter, double* history,
                  oc_stats* st) {
  memset(st, 0, sizeof *st);
  double* inv_diag = (double*)malloc((size_t)(n > 0 ? n : 1) * sizeof(double));
  if (!inv_diag) return OC_ERR_ALLOC;
  for (int i = 0; i < n; ++i) {
    const int kd = csr_find(row_ptr, cols, i, i);
    const double d = kd >= 0 ? vals[kd] : 0.0;
    if (!(d > 0.0)) {
      st->iterations = i;
      free(inv_diag);
      return OC_ERR_BAD_DIAGONAL;
    }
    inv_diag[i] = 1.0 / d;
  }
  for (int i = 0; i < n; ++i) x[i] = 0.0;
  const double bnorm = sqrt(dot(n, b, b));
  if (bnorm == 0.0) {
    st->converged = 1;
    st->zero_rhs = 1;
    if (history) history[0] = 0.0;
    free(inv_diag);
    return OC_OK;
  }
  double* r = (double*)malloc((size_t)n * sizeof(double));
  double* z = (double*)malloc((size_t)n * sizeof(double));
  double* p = (double*)malloc((size_t)n * sizeof(double));
  double* Ap = (double*)malloc((size_t)n * sizeof(double));
  if (!r || !z || !p || !Ap) {
    free(inv_diag); free(r); free(z); free(p); free(Ap);
    return OC_ERR_ALLOC;
  }
  memcpy(r, b, (size_t)n * sizeof(double));
  for (int i = 0; i < n; ++i) z[i] = inv_diag[i] * r[i];
  memcpy(p, z, (size_t)n * sizeof(double));
  double rz = dot(n, r, z);
  st->rel_residual = 1.0;
  if (history) history[0] = 1.0;
  int done = 0;
  for (int k = 1; k <= max_iter && !done; ++k) {
    oc_spmv(n, row_ptr, cols, vals, p, Ap);
    const double pAp = dot(n, p, Ap);
    if (!(pAp > 0.0)) {
      st->breakdown = 1;
      st->breakdown_pap = pAp;
      done = 2;
      break;
    }
    const double alpha = rz / pAp;
    for (int i = 0; i < n; ++i) {
      x[i] += alpha * p[i];
      r[i] -= alpha * Ap[i];
    }
    st->iterations = k;
    st->rel_residual = sqrt(dot(n, r, r)) / bnorm;
    if (history) history[k] = st->rel_residual;
    if (st->rel_residual < rtol) {
      st->converged = 1;
      done = 1;
      break;
    }
    for (int i = 0; i < n; ++i) z[i] = inv_diag[i] * r[i];
    const double rz_new = dot(n, r, z);
    const double beta = rz_new / rz;
    rz = rz_new;
    for (int i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
  }
  /* pcg_jacobi: true residual (solver.cpp:113-122) */
  oc_spmv(n, row_ptr, cols, vals, x, r);
  for (int i = 0; i < n; ++i) r[i] = b[i] - r[i];
  st->true_rel_residual = sqrt(dot(n, r, r)) / bnorm;
  free(inv_diag); free(r); free(z); free(p); free(Ap);
  return st->converged ? OC_OK : OC_NOT_CONVERGED;
}

/* Jacobi-PCG of solver.cpp:35-131 on `threads` OpenMP threads (timed baseline; see the
 * header).  Rows are split statically; each dot product sums per-thread partials in thread
 * order.  Returns as oc_pcg_jacobi (history is not recorded). */
static double dot_mt(int n, const double* a, const double* b, double* part, int threads) {
#pragma omp parallel num_threads(threads)
  {
    double s = 0.0;
#pragma omp for schedule(static)
    for (int i = 0; i < n; ++i) s += a[i] * b[i];
#ifdef _OPENMP
    part[omp_get_thread_num()] = s;
#else
    part[0] = s;
#endif
  }
  double t = 0.0;
  for (int k = 0; k < threads; ++k) t += part[k];
  return t;
}

int oc_pcg_jacobi_mt(int n, const int* row_ptr, const int* cols, const double* vals,
                     const double* b, double* x, double rtol, int max_iter, int threads,
                     oc_stats* st) {
  memset(st, 0, sizeof *st);
  if (threads < 1) threads = 1;
#ifndef _OPENMP
  threads = 1;
#endif
  double* part = (double*)calloc((size_t)threads, sizeof(double));
  double* inv_diag = (double*)malloc((size_t)(n > 0 ? n : 1) * sizeof(double));
  double* r = (double*)malloc((size_t)(n > 0 ? n : 1) * sizeof(double));
  double* p = (double*)malloc((size_t)(n > 0 ? n : 1) * sizeof(double));
  double* Ap = (double*)malloc((size_t)(n > 0 ? n : 1) * sizeof(double));
  if (!part || !inv_diag || !r || !p || !Ap) {
    free(part); free(inv_diag); free(r); free(p); free(Ap);
    return OC_ERR_ALLOC;
  }
  int bad = -1;
  for (int i = 0; i < n; ++i) {
    const int kd = csr_find(row_ptr, cols, i, i);
    const double d = kd >= 0 ? vals[kd] : 0.0;
    if (!(d > 0.0)) { bad = i; break; }
    inv_diag[i] = 1.0 / d;
  }
  int rc = OC_OK;
  if (bad >= 0) {
    st->iterations = bad;
    rc = OC_ERR_BAD_DIAGONAL;
    goto done;
  }
#pragma omp parallel for num_threads(threads) schedule(static)
  for (int i = 0; i < n; ++i) {
    x[i] = 0.0;
    r[i] = b[i];
    p[i] = inv_diag[i] * b[i];
  }
  {
    const double bnorm = sqrt(dot_mt(n, b, b, part, threads));
    if (bnorm == 0.0) {
      st->converged = 1;
      st->zero_rhs = 1;
      goto done;
    }
    double rz = dot_mt(n, r, p, part, threads);
    st->rel_residual = 1.0;
    for (int k = 1; k <= max_iter; ++k) {
#pragma omp parallel for num_threads(threads) schedule(static)
      for (int i = 0; i < n; ++i) {
        double s = 0.0;
        for (int q = row_ptr[i]; q < row_ptr[i + 1]; ++q) s += vals[q] * p[cols[q]];
        Ap[i] = s;
      }
      const double pAp = dot_mt(n, p, Ap, part, threads);
      if (!(pAp > 0.0)) {
        st->breakdown = 1;
        st->breakdown_pap = pAp;
        break;
      }
      const double alpha = rz / pAp;
#pragma omp parallel for num_threads(threads) schedule(static)
      for (int i = 0; i < n; ++i) {
        x[i] += alpha * p[i];
        r[i] -= alpha * Ap[i];
      }
      st->iterations = k;
      st->rel_residual = sqrt(dot_mt(n, r, r, part, threads)) / bnorm;
      if (st->rel_residual < rtol) {
        st->converged = 1;
        break;
      }
      /* z = D^-1 r is folded into the r'z product and the direction update */
      double rz_new = 0.0;
#pragma omp parallel num_threads(threads)
      {
        double s = 0.0;
#pragma omp for schedule(static)
        for (int i = 0; i < n; ++i) s += r[i] * (inv_diag[i] * r[i]);
#ifdef _OPENMP
        part[omp_get_thread_num()] = s;
#else
        part[0] = s;
#endif
      }
      for (int t = 0; t < threads; ++t) rz_new += part[t];
      const double beta = rz_new / rz;
      rz = rz_new;
#pragma omp parallel for num_threads(threads) schedule(static)
      for (int i = 0; i < n; ++i) p[i] = inv_diag[i] * r[i] + beta * p[i];
    }
#pragma omp parallel for num_threads(threads) schedule(static)
    for (int i = 0; i < n; ++i) {
      double s = 0.0;
      for (int q = row_ptr[i]; q < row_ptr[i + 1]; ++q) s += vals[q] * x[cols[q]];
      r[i] = b[i] - s;
    }
    st->true_rel_residual = sqrt(dot_mt(n, r, r, part, threads)) / bnorm;
    rc = st->converged ? OC_OK : OC_NOT_CONVERGED;
  }
done:
  free(part); free(inv_diag); free(r); free(p); free(Ap);
  return rc;
}

/* postprocess.cpp:35-60: constant stress of each element (sxx, syy, txy), centroid, area. */
void oc_element_stresses(int64_t ne, const int* tri, const double* xy, const double* u,
                         const double* E_elem, double E, double nu, int plane_strain,
                         double* sig /* 3*ne */, double* cxy /* 2*ne */, double* area) {
  double D[9];
  oc_constitutive(E, nu, plane_strain, D);
  for (int64_t e = 0; e < ne; ++e) {
    double exy[6], ue[6], eps[3];
    elem_coords(tri, xy, e, exy);
    for (int k = 0; k < 3; ++k) {
      const int n = tri[3 * e + k];
      ue[2 * k] = u[2 * n];
      ue[2 * k + 1] = u[2 * n + 1];
    }
    if (E_elem) oc_constitutive(E_elem[e], nu, plane_strain, D);
    oc_strain(exy, ue, eps);
    for (int i = 0; i < 3; ++i) {
      double s = 0.0;
      for (int j = 0; j < 3; ++j) s += D[i * 3 + j] * eps[j];
      sig[3 * e + i] = s;
    }
    cxy[2 * e] = (exy[0] + exy[2] + exy[4]) / 3.0;
    cxy[2 * e + 1] = (exy[1] + exy[3] + exy[5]) / 3.0;
    area[e] = fabs(oc_signed_area(exy));
  }
}

/* postprocess.cpp:81-150 (kirsch_partial + set_relative_errors + set_scf).
 * out: [l2_stress, l2_disp, energy, scf_element, scf_x, scf_y, scf_num, scf_analytic] */
void oc_kirsch_metrics(int64_t ne, const int* tri, const double* xy, const double* u,
                       const double* sig, const double* cxy, const double* area, double sigma_inf,
                       double a, double E, double nu, int plane_strain, double out[8]) {
  double D[9];
  oc_constitutive(E, nu, plane_strain, D);
  const double det = D[0] * D[4] - D[1] * D[3];
  const double C00 = D[4] / det, C01 = -D[1] / det, C11 = D[0] / det, C22 = 1.0 / D[8];
  const double bary[3][3] = {{2.0 / 3, 1.0 / 6, 1.0 / 6}, {1.0 / 6, 2.0 / 3, 1.0 / 6},
                             {1.0 / 6, 1.0 / 6, 2.0 / 3}};
  double err2 = 0, ref2 = 0, derr2 = 0, dref2 = 0, eerr2 = 0, eref2 = 0;
  double best = INFINITY;
  int64_t scf = -1;
  for (int64_t e = 0; e < ne; ++e) {
    double ex[3];
    const double sx = sig[3 * e], sy = sig[3 * e + 1], txy = sig[3 * e + 2];
    oc_kirsch_stress(sigma_inf, a, cxy[2 * e], cxy[2 * e + 1], ex);
    const double d0 = sx - ex[0], d1 = sy - ex[1], d2 = txy - ex[2];
    err2 += area[e] * (d0 * d0 + d1 * d1 + 2 * d2 * d2);
    ref2 += area[e] * (ex[0] * ex[0] + ex[1] * ex[1] + 2 * ex[2] * ex[2]);
    const double d = hypot(cxy[2 * e], cxy[2 * e + 1] - a);
    if (d < best) { best = d; scf = e; }
    const int* n = &tri[3 * e];
    const double w = area[e] / 3.0;
    for (int q = 0; q < 3; ++q) {
      double x = 0, y = 0, uh = 0, vh = 0;
      for (int i = 0; i < 3; ++i) {
        x += bary[q][i] * xy[2 * n[i]];
        y += bary[q][i] * xy[2 * n[i] + 1];
        uh += bary[q][i] * u[2 * n[i]];
        vh += bary[q][i] * u[2 * n[i] + 1];
      }
      double ua[2], sq[3];
      oc_kirsch_displacement(sigma_inf, a, E, nu, plane_strain, x, y, ua);
      derr2 += w * ((uh - ua[0]) * (uh - ua[0]) + (vh - ua[1]) * (vh - ua[1]));
      dref2 += w * (ua[0] * ua[0] + ua[1] * ua[1]);
      oc_kirsch_stress(sigma_inf, a, x, y, sq);
      const double v0 = sx - sq[0], v1 = sy - sq[1], v2 = txy - sq[2];
      eerr2 += w * (C00 * v0 * v0 + 2.0 * C01 * v0 * v1 + C11 * v1 * v1 + C22 * v2 * v2);
      eref2 += w * (C00 * sq[0] * sq[0] + 2.0 * C01 * sq[0] * sq[1] + C11 * sq[1] * sq[1] +
                    C22 * sq[2] * sq[2]);
    }
  }
  out[0] = sqrt(err2 / ref2);
  out[1] = sqrt(derr2 / dref2);
  out[2] = sqrt(eerr2 / eref2);
  out[3] = (double)scf;
  if (scf >= 0) {
    double ex[3];
    oc_kirsch_stress(sigma_inf, a, cxy[2 * scf], cxy[2 * scf + 1], ex);
    out[4] = cxy[2 * scf];
    out[5] = cxy[2 * scf + 1];
    out[6] = sig[3 * scf] / sigma_inf;
    out[7] = ex[0] / sigma_inf;
  }
}
