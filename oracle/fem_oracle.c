/*
 * fem_oracle.c -- CPU restatement of SCI-Solver_FEM's assembly + AMG-PCG path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the GPU library is
 * compared against and the CPU baseline that bench.py times; the product
 * (libfem_mi355x.so) never links or calls it.
 *
 * Every function names the reference code it restates (paths relative to the
 * SCI-Solver_FEM repository root).  Deliberate deviations are listed in the
 * function comments.  Plain C11 + optional OpenMP; no dependencies.
 */
#include "fem_oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------- */
/* Pattern: TetMesh::need_neighbors (src/core/cuda/tetmesh.cu:112-170),      */
/* TriMesh::need_neighbors (src/core/aggmis/cuda/TriMesh_connectivity.cu:     */
/* 96-131) and tetmesh2ell / trimesh2ell (src/core/cuda/cutil.cu:169-230,     */
/* 298-348): row i = {i} U {vertices sharing an element with i}.  The         */
/* reference stores [i, sorted neighbours] in ELL; the canonical CSR returned  */
/* here holds the same set in ascending order.                                */
/* ------------------------------------------------------------------------- */
typedef struct {
  int32_t* v;
  int32_t n, cap;
} ivec;

static void ivec_push(ivec* a, int32_t x) {
  if (a->n == a->cap) {
    a->cap = a->cap ? 2 * a->cap : 16;
    a->v = (int32_t*)realloc(a->v, sizeof(int32_t) * (size_t)a->cap);
  }
  a->v[a->n++] = x;
}

static int cmp_i32(const void* a, const void* b) {
  int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
  return (x > y) - (x < y);
}

int64_t orc_pattern(int npe, int64_t nv, int64_t ne, const int32_t* conn, int64_t* rp, int32_t* ci) {
  ivec* nb = (ivec*)calloc((size_t)nv, sizeof(ivec));
  for (int64_t e = 0; e < ne; ++e) {
    const int32_t* t = conn + npe * e;
    for (int j = 0; j < npe; ++j) {
      ivec* me = &nb[t[j]];
      for (int o = 1; o < npe; ++o) {
        int32_t c = t[(j + o) % npe];
        int found = 0;
        for (int32_t q = 0; q < me->n; ++q)
          if (me->v[q] == c) { found = 1; break; }
        if (!found) ivec_push(me, c);
      }
    }
  }
  int64_t nnz = 0;
  rp[0] = 0;
  for (int64_t i = 0; i < nv; ++i) {
    ivec* me = &nb[i];
    /* the row is the neighbour set plus the diagonal */
    int has_diag = 0;
    for (int32_t q = 0; q < me->n; ++q) has_diag |= (me->v[q] == (int32_t)i);
    if (!has_diag) ivec_push(me, (int32_t)i);
    qsort(me->v, (size_t)me->n, sizeof(int32_t), cmp_i32);
    if (ci) memcpy(ci + nnz, me->v, sizeof(int32_t) * (size_t)me->n);
    nnz += me->n;
    rp[i + 1] = nnz;
    free(me->v);
  }
  free(nb);
  return nnz;
}

/* ------------------------------------------------------------------------- */
/* Gauss-Jacobi quadrature (FEM3D::JacobiGZeros / JacobiGLZW / JacobiGRZW,   */
/* src/core/cuda/FEM3D.cu:138-368; FEM2D.cu:358-378).  Restated from the      */
/* textbook formulas (Karniadakis & Sherwin, Spectral/hp Element Methods,     */
/* App. B): zeros by Newton iteration with polynomial deflation from          */
/* Chebyshev-Gauss starting points, weights in closed form with Gamma.       */
/* ------------------------------------------------------------------------- */
static double jacobi_p(int n, double x, double a, double b) {
  if (n == 0) return 1.0;
  double p0 = 1.0, p1 = 0.5 * (a - b + (a + b + 2.0) * x);
  for (int k = 2; k <= n; ++k) {
    double km1 = k - 1.0;
    double t = 2.0 * km1 + a + b;
    double a1 = 2.0 * (km1 + 1.0) * (km1 + a + b + 1.0) * t;
    double a2 = (t + 1.0) * (a * a - b * b);
    double a3 = t * (t + 1.0) * (t + 2.0);
    double a4 = 2.0 * (km1 + a) * (km1 + b) * (t + 2.0);
    double p2 = ((a2 + a3 * x) * p1 - a4 * p0) / a1;
    p0 = p1;
    p1 = p2;
  }
  return p1;
}

static double jacobi_dp(int n, double x, double a, double b) {
  if (n == 0) return 0.0;
  return 0.5 * (a + b + n + 1.0) * jacobi_p(n - 1, x, a + 1.0, b + 1.0);
}

static void jacobi_zeros(int n, double a, double b, double* z) {
  const double pi = 3.14159265358979323846;
  double last = 0.0;
  for (int k = 0; k < n; ++k) {
    double r = -cos((2.0 * k + 1.0) * pi / (2.0 * n));
    if (k) r = 0.5 * (r + last);
    for (int it = 0; it < 100; ++it) {
      double s = 0.0;
      for (int i = 0; i < k; ++i) s += 1.0 / (r - z[i]);
      double p = jacobi_p(n, r, a, b);
      double dr = -p / (jacobi_dp(n, r, a, b) - s * p);
      r += dr;
      if (fabs(dr) < 1e-15) break;
    }
    z[k] = r;
    last = r;
  }
}

void orc_gauss_lobatto_jacobi(int q, double a, double b, double* z, double* w) {
  if (q == 1) {
    z[0] = 0.0;
    w[0] = 2.0;
    return;
  }
  z[0] = -1.0;
  z[q - 1] = 1.0;
  jacobi_zeros(q - 2, a + 1.0, b + 1.0, z + 1);
  double fac = pow(2.0, a + b + 1.0) * tgamma(a + q) * tgamma(b + q) /
               ((q - 1.0) * tgamma((double)q) * tgamma(a + b + q + 1.0));
  for (int i = 0; i < q; ++i) {
    double p = jacobi_p(q - 1, z[i], a, b);
    w[i] = fac / (p * p);
  }
  w[0] *= (b + 1.0);
  w[q - 1] *= (a + 1.0);
}

void orc_gauss_radau_jacobi(int q, double a, double b, double* z, double* w) {
  if (q == 1) {
    z[0] = 0.0;
    w[0] = 2.0;
    return;
  }
  z[0] = -1.0;
  jacobi_zeros(q - 1, a, b + 1.0, z + 1);
  double fac = pow(2.0, a + b) * tgamma(a + q) * tgamma(b + q) /
               (tgamma((double)q) * (b + q) * tgamma(a + b + q + 1.0));
  for (int i = 0; i < q; ++i) {
    double p = jacobi_p(q - 1, z[i], a, b);
    w[i] = fac * (1.0 - z[i]) / (p * p);
  }
  w[0] *= (b + 1.0);
}

/* FEM3D::assemble (src/core/cuda/FEM3D.cu:501-557) + IntegrationInTet
 * (:452-475): 4-point Lobatto(0,0) x Radau(1,0)/2 x Radau(2,0)/4 rule in
 * collapsed coordinates; out[10] = quadrature of phi_k phi_g (k <= g) over the
 * collapsed cube, i.e. 8 * the integral over the unit reference tet. */
void orc_tet_mass_integrals(double out[10]) {
  enum { Q = 4 };
  double zx[Q], wx[Q], zy[Q], wy[Q], zz[Q], wz[Q];
  orc_gauss_lobatto_jacobi(Q, 0, 0, zx, wx);
  orc_gauss_radau_jacobi(Q, 1, 0, zy, wy);
  orc_gauss_radau_jacobi(Q, 2, 0, zz, wz);
  for (int i = 0; i < Q; ++i) {
    wy[i] /= 2.0;
    wz[i] /= 4.0;
  }
  int cnt = 0;
  for (int k = 0; k < 4; ++k)
    for (int g = k; g < 4; ++g) {
      double integral = 0.0;
      for (int i = 0; i < Q; ++i) {
        double ty = 0.0;
        for (int j = 0; j < Q; ++j) {
          double tz = 0.0;
          for (int l = 0; l < Q; ++l) {
            double x = (1 + zx[i]) * 0.5 * (1 - zy[j]) * 0.5 * (1 - zz[l]) * 0.5;
            double y = (1 + zy[j]) * 0.5 * (1 - zz[l]) * 0.5;
            double z = (1 + zz[l]) * 0.5;
            double phi[4] = {1.0 - x - y - z, x, y, z};
            tz += phi[k] * phi[g] * wz[l];
          }
          ty += tz * wy[j];
        }
        integral += ty * wx[i];
      }
      out[cnt++] = integral;
    }
}

/* 2D counterpart (FEM2D.cu:358-378, perform_element_loop_2D.cuh:72-147):
 * 6-point Lobatto(0,0) x Radau(1,0) rule on the collapsed square (the 2D
 * weights are not rescaled); out[6] = quadrature of phi_k phi_g (k <= g),
 * i.e. 8 * the integral over the unit reference triangle. */
void orc_tri_mass_integrals(double out[6]) {
  enum { Q = 6 };
  double zx[Q], wx[Q], zy[Q], wy[Q];
  orc_gauss_lobatto_jacobi(Q, 0, 0, zx, wx);
  orc_gauss_radau_jacobi(Q, 1, 0, zy, wy);
  int cnt = 0;
  for (int k = 0; k < 3; ++k)
    for (int g = k; g < 3; ++g) {
      double integral = 0.0;
      for (int i = 0; i < Q; ++i) {
        double ty = 0.0;
        for (int j = 0; j < Q; ++j) {
          double x = (1 + zx[i]) * 0.5 * (1 - zy[j]) * 0.5;
          double y = (1 + zy[j]) * 0.5;
          double phi[3] = {1.0 - x - y, x, y};
          ty += phi[k] * phi[g] * wy[j];
        }
        integral += ty * wx[i];
      }
      out[cnt++] = integral;
    }
}

/* ------------------------------------------------------------------------- */
/* Element matrices: element_loop_3d_kernel                                  */
/* (src/core/cuda/perform_element_loop_3D.cuh:481-627, stiffness :29-48,      */
/* mass :75-116): gradients from the inverse of [x y z 1] (cofactors / det),  */
/* K_kg = (grad phi_k . grad phi_g) * Tvol * co, M_kg = integrand_kg * |det|  */
/* with det = triple product / 8.  Full symmetric 4 x 4 outputs.              */
/* ------------------------------------------------------------------------- */
static double det3(double a11, double a12, double a13, double a21, double a22, double a23, double a31, double a32,
                   double a33) {
  return a11 * (a22 * a33 - a23 * a32) - a12 * (a21 * a33 - a23 * a31) + a13 * (a21 * a32 - a22 * a31);
}

void orc_tet_element(const double* X, double co, const double integ[10], double K[16], double M[16]) {
  const double *p0 = X, *p1 = X + 3, *p2 = X + 6, *p3 = X + 9;
  /* volume as in :528-534 */
  double a1 = p1[0] - p3[0], a2 = p1[1] - p3[1], a3 = p1[2] - p3[2];
  double b1 = p2[0] - p3[0], b2 = p2[1] - p3[1], b3 = p2[2] - p3[2];
  double c1 = p0[0] - p3[0], c2 = p0[1] - p3[1], c3 = p0[2] - p3[2];
  double Tvol = fabs(c1 * (a2 * b3 - a3 * b2) + c2 * (a3 * b1 - a1 * b3) + c3 * (a1 * b2 - a2 * b1)) / 6.0;
  /* inverse of the 4x4 matrix rows (x_i, y_i, z_i, 1) by cofactors */
  double A[4][4];
  for (int i = 0; i < 4; ++i) {
    A[i][0] = X[3 * i];
    A[i][1] = X[3 * i + 1];
    A[i][2] = X[3 * i + 2];
    A[i][3] = 1.0;
  }
  double cof[4][4];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      double m[9];
      int t = 0;
      for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c)
          if (r != i && c != j) m[t++] = A[r][c];
      double d = det3(m[0], m[1], m[2], m[3], m[4], m[5], m[6], m[7], m[8]);
      cof[i][j] = ((i + j) & 1) ? -d : d;
    }
  double det = 0.0;
  for (int j = 0; j < 4; ++j) det += A[0][j] * cof[0][j];
  /* inverse = adj / det, adj = cof^T; column s of the inverse holds basis s:
   * phi_s(x) = inv[0][s] x + inv[1][s] y + inv[2][s] z + inv[3][s] */
  double g[4][3];
  for (int s = 0; s < 4; ++s)
    for (int d = 0; d < 3; ++d) g[s][d] = cof[s][d] / det;
  double jac = 0.125 * ((-p0[0] + p1[0]) * (-p0[1] + p2[1]) * (-p0[2] + p3[2]) +
                        (-p0[1] + p1[1]) * (-p0[2] + p2[2]) * (-p0[0] + p3[0]) +
                        (-p0[2] + p1[2]) * (-p0[0] + p2[0]) * (-p0[1] + p3[1]) -
                        (-p0[0] + p1[0]) * (-p0[2] + p2[2]) * (-p0[1] + p3[1]) -
                        (-p0[2] + p1[2]) * (-p0[1] + p2[1]) * (-p0[0] + p3[0]) -
                        (-p0[1] + p1[1]) * (-p0[0] + p2[0]) * (-p0[2] + p3[2]));
  jac = fabs(jac);
  int cnt = 0;
  for (int k = 0; k < 4; ++k)
    for (int q = k; q < 4; ++q) {
      double s = (g[k][0] * g[q][0] + g[k][1] * g[q][1] + g[k][2] * g[q][2]) * Tvol * co;
      double m = integ[cnt++] * jac;
      K[4 * k + q] = K[4 * q + k] = s;
      M[4 * k + q] = M[4 * q + k] = m;
    }
}

/* element_loop_kernel (src/core/cuda/perform_element_loop_2D.cuh:288-372):
 * 3x3 inverse of rows (x_i, y_i, 1); stiffness * TArea (unsigned); mass with
 * the signed Jacobian (x0 y1 - x1 y0 - x0 y2 + x2 y0 + x1 y2 - x2 y1)/8 (:92)
 * when signed_mass != 0, else its absolute value. */
void orc_tri_element(const double* X, int signed_mass, const double integ[6], double K[9], double M[9]) {
  double x[3] = {X[0], X[3], X[6]}, y[3] = {X[1], X[4], X[7]};
  double TArea = fabs(x[0] * y[2] - x[0] * y[1] + x[1] * y[0] - x[1] * y[2] + x[2] * y[1] - x[2] * y[0]) / 2.0;
  double det = x[0] * y[1] + x[1] * y[2] + x[2] * y[0] - x[0] * y[2] - x[2] * y[1] - x[1] * y[0];
  /* inverse columns: phi_s = c0 x + c1 y + c2 */
  double g[3][2];
  g[0][0] = (y[1] - y[2]) / det;
  g[0][1] = (x[2] - x[1]) / det;
  g[1][0] = (y[2] - y[0]) / det;
  g[1][1] = (x[0] - x[2]) / det;
  g[2][0] = (y[0] - y[1]) / det;
  g[2][1] = (x[1] - x[0]) / det;
  double jac = (x[0] * y[1] - x[1] * y[0] - x[0] * y[2] + x[2] * y[0] + x[1] * y[2] - x[2] * y[1]) / 8.0;
  if (!signed_mass) jac = fabs(jac);
  /* the reference integrates the physical basis over the collapsed square
   * times this Jacobian: M_kg = integ_kg * jac */
  int cnt = 0;
  for (int k = 0; k < 3; ++k)
    for (int q = k; q < 3; ++q) {
      double s = (g[k][0] * g[q][0] + g[k][1] * g[q][1]) * TArea;
      double m = integ[cnt++] * jac;
      K[3 * k + q] = K[3 * q + k] = s;
      M[3 * k + q] = M[3 * q + k] = m;
    }
}

static int64_t find_slot(const int64_t* rp, const int32_t* ci, int64_t row, int32_t col) {
  int64_t lo = rp[row], hi = rp[row + 1];
  while (lo < hi) {
    int64_t mid = (lo + hi) >> 1;
    if (ci[mid] < col) lo = mid + 1; else hi = mid;
  }
  return (lo < rp[row + 1] && ci[lo] == col) ? lo : -1;
}

/* perform_element_loop_3d / element_loop_kernel + sum_into_global_linear_
 * system (perform_element_loop_3D.cuh:164-318,758-857): element contributions
 * kappa_e * K_e + mass_coeff * M_e added into the CSR pattern, elements in
 * index order.  (The reference fixes mass_coeff = 1 and takes kappa from the
 * material-label table.) */
int orc_assemble(int dim, int64_t nv, const double* xyz, int64_t ne, const int32_t* conn, const double* kappa,
                 double mass_coeff, int signed_mass, const int64_t* rp, const int32_t* ci, double* vals) {
  const int npe = dim + 1;
  double integ3[10], integ2[6];
  orc_tet_mass_integrals(integ3);
  orc_tri_mass_integrals(integ2);
  memset(vals, 0, sizeof(double) * (size_t)rp[nv]);
  for (int64_t e = 0; e < ne; ++e) {
    const int32_t* t = conn + npe * e;
    double X[12], K[16], M[16];
    for (int a = 0; a < npe; ++a)
      for (int d = 0; d < 3; ++d) X[3 * a + d] = xyz[3 * t[a] + d];
    double co = kappa ? kappa[e] : 1.0;
    if (dim == 3) {
      orc_tet_element(X, co, integ3, K, M);
    } else {
      orc_tri_element(X, signed_mass, integ2, K, M);
      for (int q = 0; q < 9; ++q) K[q] *= co;
    }
    for (int a = 0; a < npe; ++a)
      for (int b = 0; b < npe; ++b) {
        int64_t s = find_slot(rp, ci, t[a], t[b]);
        if (s < 0) return -1;
        vals[s] += K[npe * a + b] + mass_coeff * M[npe * a + b];
      }
  }
  return 0;
}

/* ------------------------------------------------------------------------- */
/* element gradient of the adjoint: -|T| grad(a) . grad(u) (no reference     */
/* counterpart); gradients as in orc_tet_element / orc_tri_element.           */
/* ------------------------------------------------------------------------- */
void orc_element_grad(int dim, const double* xyz, int64_t ne, const int32_t* conn, const double* a, const double* u,
                      double* out) {
  const int npe = dim + 1;
  double integ3[10], integ2[6];
  orc_tet_mass_integrals(integ3);
  orc_tri_mass_integrals(integ2);
  for (int64_t e = 0; e < ne; ++e) {
    const int32_t* t = conn + npe * e;
    double X[12], K[16], M[16];
    for (int q = 0; q < npe; ++q)
      for (int d = 0; d < 3; ++d) X[3 * q + d] = xyz[3 * t[q] + d];
    /* a^T K_e(kappa=1) u = |T| grad(a).grad(u) */
    if (dim == 3) orc_tet_element(X, 1.0, integ3, K, M);
    else orc_tri_element(X, 0, integ2, K, M);
    double s = 0.0;
    for (int p = 0; p < npe; ++p)
      for (int q = 0; q < npe; ++q) s += a[t[p]] * K[npe * p + q] * u[t[q]];
    out[e] = -s;
  }
}

/* ========================================================================= */
/* Aggregation: randomized distance-k MIS + node allocation                  */
/* ========================================================================= */

static uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

/* seed of generator thread i; replaces rand() after srand(time(NULL))
 * (src/core/cuda/randomizedMIS_GPU.cu:195-198) with a seeded stream */
uint32_t orc_mis_seed(uint64_t seed, uint32_t i) {
  return (uint32_t)(mix64(seed * 0x9E3779B97F4A7C15ull + (uint64_t)(i + 1) * 0xD1B54A32D192ED03ull) >> 32);
}

static uint32_t taus_step(uint32_t z) {
  uint32_t b = (((z << 13) ^ z) >> 19);
  return ((z & 0xffffffffu) << 12) ^ b;
}

#define MIS_THREADS 32768
#define MIS_IN 1000001

/* misHelpers::randomizedMIS (src/core/cuda/randomizedMIS_GPU.cu:154-272) with
 * Generate_Randoms_Kernel (:3-20), First_Initialize_Kernel (:22-34),
 * Initialize_Kernel (:36-62), Iterate_Kernel (:64-99), Final_Iterate_Kernel
 * (:101-152).  All hops are Jacobi updates (read old arrays, write new). */
void orc_mis(int64_t n, const int32_t* ap, const int32_t* aj, int depth, uint64_t seed, int32_t* mis) {
  uint32_t* rnd = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(n ? n : 1));
  int32_t *bi = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n ? n : 1));
  int32_t *bo = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n ? n : 1));
  int32_t *oi = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n ? n : 1));
  int32_t *oo = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n ? n : 1));
  if (depth < 1) depth = 1;
  for (int64_t i = 0; i < n; ++i) mis[i] = -1;
  for (int t = 0; t < MIS_THREADS; ++t) {
    uint32_t z = orc_mis_seed(seed, (uint32_t)t);
    for (int64_t off = t; off < n; off += MIS_THREADS) {
      z = taus_step(z);
      rnd[off] = z;
    }
  }
  for (int64_t i = 0; i < n; ++i) {
    oi[i] = (int32_t)i;
    bi[i] = (int32_t)(rnd[i] % 1000000u);
  }
  int first = 1;
  for (;;) {
    if (!first) {
      for (int64_t i = 0; i < n; ++i) {
        uint32_t z = taus_step(rnd[i]);
        oi[i] = (int32_t)i;
        bi[i] = mis[i] == -1 ? (int32_t)(z % 1000000u) : (mis[i] == 1 ? MIS_IN : 0);
        rnd[i] = z;
      }
    }
    first = 0;
    int incomplete = 0;
    for (int h = 0; h < depth; ++h) {
      const int final_hop = (h == depth - 1);
#pragma omp parallel for schedule(static) reduction(| : incomplete)
      for (int64_t i = 0; i < n; ++i) {
        int32_t best = bi[i], org = oi[i];
        if (best < MIS_IN) {
          for (int32_t q = ap[i]; q < ap[i + 1]; ++q) {
            int32_t nb = aj[q], ch = bi[nb], cho = oi[nb];
            if (ch > 0 && ch == best && cho > org) org = cho;
            if (ch > best) {
              best = ch;
              org = cho;
            }
          }
        }
        if (!final_hop) {
          bo[i] = best;
          oo[i] = org;
        } else {
          int32_t st = -1;
          if (org == (int32_t)i) st = 1;
          else if (best == MIS_IN) st = 0;
          mis[i] = st;
          if (st == -1) incomplete = 1;
        }
      }
      if (!final_hop) {
        int32_t* t = bi; bi = bo; bo = t;
        t = oi; oi = oo; oo = t;
      }
    }
    if (!incomplete) break;
  }
  free(rnd); free(bi); free(bo); free(oi); free(oo);
}

/* misHelpers::aggregateGraph (src/core/cuda/misHelpers.cu:513-601) with
 * allocateNodesKernel (:13-75) and removeRuntyParts (:744-781).
 * rule 0: a node joins the aggregate of its first aggregated neighbour in
 *   adjacency order -- what the reference kernel computes: its candidate
 *   table fills all ten slots with the first candidate, so later candidates
 *   are never counted;
 * rule 1: the most frequent neighbouring aggregate, ties to the smallest id.
 * Deviation: nodes that no aggregate can reach (a component emptied by runt
 * removal) become singleton aggregates and the loop ends; the reference
 * loops forever there.  Returns the number of aggregates. */
int64_t orc_aggregate(int64_t n, const int32_t* ap, const int32_t* aj, int depth, int min_size, int rule,
                      uint64_t seed, int32_t* agg) {
  if (n == 0) return 0;
  int32_t* mis = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
  int32_t* pin = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
  int32_t* pout = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
  int32_t* aggd = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
  orc_mis(n, ap, aj, depth, seed, mis);
  int32_t cnt = 0;
  for (int64_t i = 0; i < n; ++i) {
    if (mis[i] == 1) {
      pin[i] = cnt++;
      aggd[i] = 1;
    } else {
      pin[i] = -1;
      aggd[i] = 0;
    }
  }
  int64_t nagg = 0;
  int64_t prev_unalloc = -1;
  for (;;) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
      if (aggd[i]) {
        pout[i] = pin[i];
        continue;
      }
      int32_t add = -1;
      if (rule == 0) {
        for (int32_t q = ap[i]; q < ap[i + 1]; ++q)
          if (pin[aj[q]] != -1) {
            add = pin[aj[q]];
            break;
          }
      } else {
        int best = 0;
        for (int32_t q = ap[i]; q < ap[i + 1]; ++q) {
          int32_t c = pin[aj[q]];
          if (c == -1) continue;
          int seen = 0;
          for (int32_t r = ap[i]; r < q; ++r)
            if (pin[aj[r]] == c) { seen = 1; break; }
          if (seen) continue;
          int k = 0;
          for (int32_t r = q; r < ap[i + 1]; ++r) k += (pin[aj[r]] == c);
          if (k > best || (k == best && c < add)) {
            best = k;
            add = c;
          }
        }
      }
      pout[i] = add;
      if (add != -1) aggd[i] = 1;
    }
    int32_t* t = pin; pin = pout; pout = t;
    int64_t unalloc = 0;
    for (int64_t i = 0; i < n; ++i) unalloc += (aggd[i] == 0);
    if (unalloc == 0) {
      int32_t mx = -1;
      for (int64_t i = 0; i < n; ++i) if (pin[i] > mx) mx = pin[i];
      int64_t np = mx + 1;
      int32_t* sz = (int32_t*)calloc((size_t)np, sizeof(int32_t));
      int32_t* ex = (int32_t*)malloc(sizeof(int32_t) * (size_t)np);
      for (int64_t i = 0; i < n; ++i) sz[pin[i]]++;
      int32_t removed = 0;
      for (int64_t p = 0; p < np; ++p) {
        ex[p] = removed;
        removed += (sz[p] < min_size);
      }
      if (removed == 0) {
        nagg = np;
        free(sz); free(ex);
        break;
      }
      for (int64_t i = 0; i < n; ++i) {
        int32_t p = pin[i];
        if (sz[p] < min_size) {
          pin[i] = -1;
          aggd[i] = 0;
        } else {
          pin[i] = p - ex[p];
          aggd[i] = 1;
        }
      }
      free(sz); free(ex);
      prev_unalloc = -1;
      continue;
    }
    if (unalloc == prev_unalloc) {
      int32_t mx = -1;
      for (int64_t i = 0; i < n; ++i) if (pin[i] > mx) mx = pin[i];
      int32_t next = mx + 1;
      for (int64_t i = 0; i < n; ++i) if (!aggd[i]) pin[i] = next++;
      nagg = next;
      break;
    }
    prev_unalloc = unalloc;
  }
  memcpy(agg, pin, sizeof(int32_t) * (size_t)n);
  free(mis); free(pin); free(pout); free(aggd);
  return nagg;
}
