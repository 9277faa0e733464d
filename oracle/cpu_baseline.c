/*
 * CPU baseline of the CSEM forward hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain C + OpenMP restatement of the reference algorithm, run on all host
 * cores beside the GPU path in bench.py (its cpu_baseline leg) and checked
 * against the numpy oracle in tests/test_oracle_cpu_baseline.py.  Nothing in
 * the petgem_amd package links or calls it.
 *
 *  pgcpu_assemble   element matrices in the quadrature form of
 *                   computeElementalMatrices (petgem/hvfem.py:258-314):
 *                   per quadrature point N = J^-1 N_ref, curl = J^T c / det J
 *                   (J rows x_k - x_0, signed det, hvfem.py:101-119,265),
 *                   M += w N^T diag(sh, sh, sv) N, K += w curl . curl, both
 *                   times det J; then A_e = K_e - i omega mu0 M_e added into
 *                   the CSR pattern (petgem/solver.py:191-235).  The reference
 *                   basis comes tabulated per local DOF and orientation value
 *                   at the quadrature points (from the oracle's basis.py).
 *  pgcpu_dirichlet  zero rows and columns of the PEC DOFs, unit diagonal,
 *                   b = 0 there (petgem/solver.py:552-577).
 *  pgcpu_cocg       COCG with Jacobi preconditioning, zero initial guess,
 *                   stopping on the recursive residual and confirming with
 *                   the true residual ||b - A x|| / ||b|| (the KSP solve of
 *                   petgem/solver.py:584-590; the reference uses PETSc).
 *
 * Complex values are interleaved (re, im) doubles.  Build: oracle/Makefile.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <omp.h>

#define MU0 (4.0 * 3.14159265358979323846 * 1e-7)

int pgcpu_threads(void) { return omp_get_max_threads(); }

static int64_t find_slot(const int64_t *rowptr, const int32_t *cols, int64_t r, int32_t c) {
  int64_t lo = rowptr[r], hi = rowptr[r + 1] - 1;
  while (lo <= hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (cols[mid] == c) return mid;
    if (cols[mid] < c) lo = mid + 1;
    else hi = mid - 1;
  }
  return -1;
}

/*
 * coords (E,4,3), sigma (E,2), orient (E,n): orientation value of each local
 * DOF (edge flag, face code or 0), S and C (n, no, nq, 3): reference basis
 * and curls of local DOF i for orientation value o at quadrature point q,
 * w (nq); dofs (E,n); CSR pattern rowptr (ndof+1), cols; values (nnz,2) out.
 * Returns 0, or -1 when a DOF pair is missing from the pattern.
 */
int pgcpu_assemble(int64_t E, int n, int nq, int no, const double *coords, const double *sigma,
                   const int32_t *orient, const double *S, const double *C, const double *w,
                   const int32_t *dofs, int64_t ndof, const int64_t *rowptr, const int32_t *cols,
                   double omega, double *values) {
  const int64_t nnz = rowptr[ndof];
  memset(values, 0, (size_t)nnz * 2 * sizeof(double));
  const double om = omega * MU0;
  int bad = 0;
#pragma omp parallel reduction(| : bad)
  {
    double *M = malloc(sizeof(double) * n * n), *K = malloc(sizeof(double) * n * n);
    double *Nr = malloc(sizeof(double) * n * 3), *Cr = malloc(sizeof(double) * n * 3);
#pragma omp for schedule(static)
    for (int64_t e = 0; e < E; ++e) {
      const double *x = coords + 12 * e;
      double J[3][3], iJ[3][3];
      for (int r = 0; r < 3; ++r)
        for (int k = 0; k < 3; ++k) J[r][k] = x[3 * (r + 1) + k] - x[k];
      const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
      const double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
      const double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
      const double det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
      const double id = 1.0 / det;
      iJ[0][0] = c00 * id;
      iJ[1][0] = c01 * id;
      iJ[2][0] = c02 * id;
      iJ[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
      iJ[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
      iJ[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
      iJ[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
      iJ[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
      iJ[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
      const double er[3] = {sigma[2 * e], sigma[2 * e], sigma[2 * e + 1]};
      memset(M, 0, sizeof(double) * n * n);
      memset(K, 0, sizeof(double) * n * n);
      const int32_t *oe = orient + (int64_t)e * n;
      for (int q = 0; q < nq; ++q) {
        for (int i = 0; i < n; ++i) {
          const double *s = S + (((int64_t)i * no + oe[i]) * nq + q) * 3;
          const double *c = C + (((int64_t)i * no + oe[i]) * nq + q) * 3;
          for (int a = 0; a < 3; ++a) {
            Nr[3 * i + a] = iJ[a][0] * s[0] + iJ[a][1] * s[1] + iJ[a][2] * s[2];
            Cr[3 * i + a] = (c[0] * J[0][a] + c[1] * J[1][a] + c[2] * J[2][a]) / det;
          }
        }
        const double wq = w[q];
        for (int i = 0; i < n; ++i)
          for (int j = 0; j < n; ++j) {
            M[i * n + j] += wq * (Nr[3 * i] * er[0] * Nr[3 * j] + Nr[3 * i + 1] * er[1] * Nr[3 * j + 1] +
                                  Nr[3 * i + 2] * er[2] * Nr[3 * j + 2]);
            K[i * n + j] += wq * (Cr[3 * i] * Cr[3 * j] + Cr[3 * i + 1] * Cr[3 * j + 1] +
                                  Cr[3 * i + 2] * Cr[3 * j + 2]);
          }
      }
      const int32_t *de = dofs + (int64_t)e * n;
      for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
          const int64_t sl = find_slot(rowptr, cols, de[i], de[j]);
          if (sl < 0) {
            bad = 1;
            continue;
          }
          const double re = K[i * n + j] * det, im = -(om * (M[i * n + j] * det));
#pragma omp atomic
          values[2 * sl] += re;
#pragma omp atomic
          values[2 * sl + 1] += im;
        }
    }
    free(M);
    free(K);
    free(Nr);
    free(Cr);
  }
  return bad ? -1 : 0;
}

void pgcpu_dirichlet(int64_t ndof, const int64_t *rowptr, const int32_t *cols,
                     const uint8_t *mask, double *values, double *b) {
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < ndof; ++r) {
    for (int64_t k = rowptr[r]; k < rowptr[r + 1]; ++k) {
      const int32_t c = cols[k];
      if (mask[r] || mask[c]) {
        values[2 * k] = (mask[r] && c == r) ? 1.0 : 0.0;
        values[2 * k + 1] = 0.0;
      }
    }
    if (mask[r]) b[2 * r] = b[2 * r + 1] = 0.0;
  }
}

static void spmv(int64_t n, const int64_t *rowptr, const int32_t *cols, const double *v,
                 const double *x, double *y) {
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < n; ++r) {
    double re = 0.0, im = 0.0;
    for (int64_t k = rowptr[r]; k < rowptr[r + 1]; ++k) {
      const double ar = v[2 * k], ai = v[2 * k + 1];
      const double xr = x[2 * cols[k]], xi = x[2 * cols[k] + 1];
      re += ar * xr - ai * xi;
      im += ar * xi + ai * xr;
    }
    y[2 * r] = re;
    y[2 * r + 1] = im;
  }
}

/* x (ndof complex) in: ignored (zero initial guess); out: the solution.
 * Returns the iteration count; *relres = true relative residual. */
int pgcpu_cocg(int64_t n, const int64_t *rowptr, const int32_t *cols, const double *v,
               const double *b, double *x, double tol, int maxit, double *relres) {
  double *r = malloc(sizeof(double) * 2 * n), *z = malloc(sizeof(double) * 2 * n);
  double *p = malloc(sizeof(double) * 2 * n), *q = malloc(sizeof(double) * 2 * n);
  double *dinv = malloc(sizeof(double) * 2 * n);
  double bn = 0.0, rho_r = 0.0, rho_i = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : bn, rho_r, rho_i)
  for (int64_t i = 0; i < n; ++i) {
    double dr = 0.0, di = 0.0;
    for (int64_t k = rowptr[i]; k < rowptr[i + 1]; ++k)
      if (cols[k] == i) {
        dr = v[2 * k];
        di = v[2 * k + 1];
      }
    const double d2 = dr * dr + di * di;
    dinv[2 * i] = dr / d2;
    dinv[2 * i + 1] = -di / d2;
    x[2 * i] = x[2 * i + 1] = 0.0;
    r[2 * i] = b[2 * i];
    r[2 * i + 1] = b[2 * i + 1];
    z[2 * i] = dinv[2 * i] * r[2 * i] - dinv[2 * i + 1] * r[2 * i + 1];
    z[2 * i + 1] = dinv[2 * i] * r[2 * i + 1] + dinv[2 * i + 1] * r[2 * i];
    p[2 * i] = z[2 * i];
    p[2 * i + 1] = z[2 * i + 1];
    bn += b[2 * i] * b[2 * i] + b[2 * i + 1] * b[2 * i + 1];
    rho_r += r[2 * i] * z[2 * i] - r[2 * i + 1] * z[2 * i + 1];
    rho_i += r[2 * i] * z[2 * i + 1] + r[2 * i + 1] * z[2 * i];
  }
  const double tol2 = tol * tol * bn;
  int it = 0;
  double res2 = bn;
  while (it < maxit && res2 > tol2) {
    spmv(n, rowptr, cols, v, p, q);
    double mu_r = 0.0, mu_i = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : mu_r, mu_i)
    for (int64_t i = 0; i < n; ++i) {
      mu_r += p[2 * i] * q[2 * i] - p[2 * i + 1] * q[2 * i + 1];
      mu_i += p[2 * i] * q[2 * i + 1] + p[2 * i + 1] * q[2 * i];
    }
    const double den = mu_r * mu_r + mu_i * mu_i;
    const double al_r = (rho_r * mu_r + rho_i * mu_i) / den;
    const double al_i = (rho_i * mu_r - rho_r * mu_i) / den;
    double nr_r = 0.0, nr_i = 0.0, rr = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : nr_r, nr_i, rr)
    for (int64_t i = 0; i < n; ++i) {
      x[2 * i] += al_r * p[2 * i] - al_i * p[2 * i + 1];
      x[2 * i + 1] += al_r * p[2 * i + 1] + al_i * p[2 * i];
      r[2 * i] -= al_r * q[2 * i] - al_i * q[2 * i + 1];
      r[2 * i + 1] -= al_r * q[2 * i + 1] + al_i * q[2 * i];
      z[2 * i] = dinv[2 * i] * r[2 * i] - dinv[2 * i + 1] * r[2 * i + 1];
      z[2 * i + 1] = dinv[2 * i] * r[2 * i + 1] + dinv[2 * i + 1] * r[2 * i];
      nr_r += r[2 * i] * z[2 * i] - r[2 * i + 1] * z[2 * i + 1];
      nr_i += r[2 * i] * z[2 * i + 1] + r[2 * i + 1] * z[2 * i];
      rr += r[2 * i] * r[2 * i] + r[2 * i + 1] * r[2 * i + 1];
    }
    const double d2 = rho_r * rho_r + rho_i * rho_i;
    const double be_r = (nr_r * rho_r + nr_i * rho_i) / d2;
    const double be_i = (nr_i * rho_r - nr_r * rho_i) / d2;
    rho_r = nr_r;
    rho_i = nr_i;
    res2 = rr;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
      const double pr = p[2 * i], pi = p[2 * i + 1];
      p[2 * i] = z[2 * i] + be_r * pr - be_i * pi;
      p[2 * i + 1] = z[2 * i + 1] + be_r * pi + be_i * pr;
    }
    ++it;
  }
  spmv(n, rowptr, cols, v, x, q);
  double tr = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : tr)
  for (int64_t i = 0; i < n; ++i) {
    const double a = b[2 * i] - q[2 * i], c = b[2 * i + 1] - q[2 * i + 1];
    tr += a * a + c * c;
  }
  *relres = bn > 0.0 ? sqrt(tr / bn) : 0.0;
  free(r);
  free(z);
  free(p);
  free(q);
  free(dinv);
  return it;
}
