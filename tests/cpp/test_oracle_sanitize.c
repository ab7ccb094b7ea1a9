/* Every function of the C oracle (oracle/cst_oracle.c) on a small jittered plate, for the
 * AddressSanitizer/UndefinedBehaviorSanitizer build: pattern, assembly with per-element E,
 * constant and Kirsch tractions, body force, Dirichlet elimination, Jacobi-PCG (sequential
 * and threaded), stresses and the Kirsch metrics; plus the error paths (bad Dirichlet DOF,
 * missing traction edge).  Exit status 0 when the solves converge and the errors are seen. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

typedef struct {
  int iterations, converged, breakdown, zero_rhs;
  double rel_residual, true_rel_residual, breakdown_pap;
} oc_stats;

int64_t oc_build_pattern(int nn, int64_t ne, const int* tri, int* row_ptr, int* cols);
int oc_assemble(int nn, int64_t ne, const int* tri, const double* xy, const double* E_elem,
                double E, double nu, int plane_strain, double t, const int* row_ptr,
                const int* cols, double* vals, int* bad_row, int* bad_col);
int oc_apply_dirichlet(int n, const int* row_ptr, const int* cols, double* vals, double* b,
                       int m, const int* dofs, const double* values, int* bad);
int oc_add_traction(int nn, int64_t ne, const int* tri, const double* xy, int nedge,
                    const int* edges, double t, int mode, double p0, double p1, double* f,
                    int* bad_edge);
void oc_add_body_force(int64_t ne, const int* tri, const double* xy, double t, double bx,
                       double by, double* f);
int oc_pcg_jacobi(int n, const int* row_ptr, const int* cols, const double* vals,
                  const double* b, double* x, double rtol, int max_iter, double* history,
                  oc_stats* st);
int oc_pcg_jacobi_mt(int n, const int* row_ptr, const int* cols, const double* vals,
                     const double* b, double* x, double rtol, int max_iter, int threads,
                     oc_stats* st);
void oc_element_stresses(int64_t ne, const int* tri, const double* xy, const double* u,
                         const double* E_elem, double E, double nu, int plane_strain,
                         double* sig, double* cxy, double* area);
void oc_kirsch_metrics(int64_t ne, const int* tri, const double* xy, const double* u,
                       const double* sig, const double* cxy, const double* area, double sigma_inf,
                       double a, double E, double nu, int plane_strain, double out[8]);

int main(void) {
  const int nx = 24, ny = 16, nn = (nx + 1) * (ny + 1), n = 2 * nn;
  const int64_t ne = 2 * (int64_t)nx * ny;
  double* xy = malloc(sizeof(double) * 2 * nn);
  int* tri = malloc(sizeof(int) * 3 * ne);
  double* E = malloc(sizeof(double) * ne);
  srand(3);
  for (int j = 0; j <= ny; ++j)
    for (int i = 0; i <= nx; ++i) {
      const int k = j * (nx + 1) + i;
      const int inner = i > 0 && i < nx && j > 0 && j < ny;
      const double h = 1.0 / nx;
      xy[2 * k] = 2.0 * i / nx + (inner ? 0.2 * h * (2.0 * rand() / RAND_MAX - 1.0) : 0.0);
      xy[2 * k + 1] = 1.0 * j / ny + (inner ? 0.2 * h * (2.0 * rand() / RAND_MAX - 1.0) : 0.0);
    }
  for (int j = 0, e = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i, e += 2) {
      const int a = j * (nx + 1) + i, b = a + 1, c = a + nx + 2, d = a + nx + 1;
      int* t = tri + 3 * e;
      t[0] = a; t[1] = b; t[2] = c; t[3] = a; t[4] = c; t[5] = d;
      E[e] = exp(0.5 * (2.0 * rand() / RAND_MAX - 1.0));
      E[e + 1] = E[e];
    }
  int* rp = malloc(sizeof(int) * (n + 1));
  int* cols = malloc(sizeof(int) * 18 * n);
  const int64_t nnz = oc_build_pattern(nn, ne, tri, rp, cols);
  double* vals = malloc(sizeof(double) * nnz);
  int br = -1, bc = -1;
  int fails = 0;
  if (oc_assemble(nn, ne, tri, xy, E, 1.0, 0.3, 1, 1.0, rp, cols, vals, &br, &bc) != 0) ++fails;
  /* right edge: constant traction, then the Kirsch traction on the same edges */
  int* edges = malloc(sizeof(int) * 2 * ny);
  for (int j = 0; j < ny; ++j) {
    edges[2 * j] = j * (nx + 1) + nx;
    edges[2 * j + 1] = (j + 1) * (nx + 1) + nx;
  }
  double* f = calloc((size_t)n, sizeof(double));
  int bad = -1;
  if (oc_add_traction(nn, ne, tri, xy, ny, edges, 1.0, 0, 1.0, 0.0, f, &bad) != 0) ++fails;
  if (oc_add_traction(nn, ne, tri, xy, ny, edges, 1.0, 1, 1.0, 0.1, f, &bad) != 0) ++fails;
  oc_add_body_force(ne, tri, xy, 1.0, 0.0, -1.0, f);
  int wrong[2] = {0, nx + 2}; /* not a boundary edge: the traction must fail */
  if (oc_add_traction(nn, ne, tri, xy, 1, wrong, 1.0, 0, 1.0, 0.0, f, &bad) == 0) ++fails;
  /* ux = 0 on x = 0, uy = 0 on y = 0 */
  int* dofs = malloc(sizeof(int) * (ny + 1 + nx + 1));
  double* dv = calloc((size_t)(ny + 1 + nx + 1), sizeof(double));
  int m = 0;
  for (int j = 0; j <= ny; ++j) dofs[m++] = 2 * (j * (nx + 1));
  for (int i = 0; i <= nx; ++i) dofs[m++] = 2 * i + 1;
  double* vcopy = malloc(sizeof(double) * nnz);
  double* fcopy = malloc(sizeof(double) * n);
  for (int64_t q = 0; q < nnz; ++q) vcopy[q] = vals[q];
  for (int i = 0; i < n; ++i) fcopy[i] = f[i];
  if (oc_apply_dirichlet(n, rp, cols, vals, f, m, dofs, dv, &bad) != 0) ++fails;
  int out_of_range = n;
  if (oc_apply_dirichlet(n, rp, cols, vcopy, fcopy, 1, &out_of_range, dv, &bad) == 0) ++fails;
  double* x = calloc((size_t)n, sizeof(double));
  double* hist = malloc(sizeof(double) * 2001);
  oc_stats st, st2;
  oc_pcg_jacobi(n, rp, cols, vals, f, x, 1e-10, 2000, hist, &st);
  double* x2 = calloc((size_t)n, sizeof(double));
  oc_pcg_jacobi_mt(n, rp, cols, vals, f, x2, 1e-10, 2000, 4, &st2);
  if (!st.converged || !st2.converged || abs(st.iterations - st2.iterations) > 2) ++fails;
  double* sig = malloc(sizeof(double) * 3 * ne);
  double* cxy = malloc(sizeof(double) * 2 * ne);
  double* area = malloc(sizeof(double) * ne);
  oc_element_stresses(ne, tri, xy, x, E, 1.0, 0.3, 1, sig, cxy, area);
  double metrics[8];
  oc_kirsch_metrics(ne, tri, xy, x, sig, cxy, area, 1.0, 0.1, 1.0, 0.3, 1, metrics);
  printf("%d DOFs, %lld entries: %d iterations (threaded %d), true rel %.3e, max sxx %.4f\n",
         n, (long long)nnz, st.iterations, st2.iterations, st.true_rel_residual, metrics[0]);
  free(xy); free(tri); free(E); free(rp); free(cols); free(vals); free(edges); free(f);
  free(dofs); free(dv); free(vcopy); free(fcopy); free(x); free(hist); free(x2); free(sig);
  free(cxy); free(area);
  return fails;
}
