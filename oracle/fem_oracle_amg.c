/*
 * fem_oracle_amg.c -- the reference AMG-PCG restated on host cores (CPU
 * baseline).  TEST INFRASTRUCTURE ONLY: timed by bench.py as the CPU
 * baseline and used by tests as a second solver; never linked into the
 * product.
 *
 * Restates, with OpenMP:
 *   AMG::setup (src/core/cuda/amg.cu:79-144): levels until < topSize rows,
 *     coarsest solved by a dense LU in the hierarchy precision (float);
 *   SmoothedMG_AMG_Level::createNextLevel (smoothedMG_amg_level.cu:403-494):
 *     randomized MIS-2 aggregation with min size 9 (CP::OldMIS,
 *     ComputePermutationMethods.cu:22-150), partitions of aggregates by a
 *     weighted MIS-1 on the induced graph (aggregateWeightedGraph,
 *     misHelpers.cu:603-677), P = (I - omega D^-1 A) T (:321-387), R = P^T,
 *     A_c = R (A P) (:390-400), all in float (AMGType, types.h:9);
 *   the partition-local weighted Jacobi smoother (preRRSym_kernel1 /
 *     preAout_kernel / postRelaxSym_kernel1, gauss_seidel.cu:1279-1376,
 *     1978-2015, 3627-3736): pre: x = w b/d, nu_pre sweeps on the
 *     in-partition block, r = b - A x; post: x += P x_c, then postRelaxes
 *     times [b_loc = b - A_out x; nu_post in-partition sweeps];
 *   AMG_Level::cycle (amg_level.cu:22-71) and CG_Flex_Cycle (cgcycle.cu:7-69).
 * Deviations (each keeps the algorithm and fixes a defect or replaces an
 * unbuildable dependency): seeded MIS instead of time(); no level permutation
 * (index bookkeeping instead, so x stays in user order); the outer CG uses the
 * FP64 operator and keeps its FP64 residual (the reference rounds both to
 * float, which floors the true residual near 1e-7); convergence is re-checked
 * on the true residual; oversized partitions (> partMax vertices) are split
 * in aggregate order instead of the swap heuristic of restrictPartitionSize
 * (misHelpers.cu:679-727).
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "fem_oracle.h"

#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
  int32_t* v;
  int32_t n, cap;
} avec;

static void avec_push(avec* a, int32_t x) {
  if (a->n == a->cap) {
    a->cap = a->cap ? 2 * a->cap : 16;
    a->v = (int32_t*)realloc(a->v, sizeof(int32_t) * (size_t)a->cap);
  }
  a->v[a->n++] = x;
}

static int cmp_int32(const void* a, const void* b) {
  int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
  return (x > y) - (x < y);
}

typedef struct {
  int64_t n, m, nnz;
  int64_t* rp;
  int32_t* ci;
  float* v;
} fcsr;

typedef struct {
  fcsr A, P, R;
  float* diag;
  int32_t* agg;
  int64_t nc;
  int64_t nparts;
  int64_t* pptr; /* partition -> rows (prow) */
  int32_t* prow;
  int32_t* part_of; /* row -> partition */
  float *x, *b, *r, *y, *xo;
} olevel;

static double now_s(void) {
#ifdef _OPENMP
  return omp_get_wtime();
#else
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
#endif
}

static void fcsr_free(fcsr* a) {
  free(a->rp);
  free(a->ci);
  free(a->v);
  memset(a, 0, sizeof(*a));
}

/* C = A * B (float, row-wise Gustavson with a dense accumulator per thread) */
static void fcsr_spgemm(const fcsr* A, const fcsr* B, fcsr* C) {
  const int64_t n = A->n, m = B->m;
  int64_t* cnt = (int64_t*)calloc((size_t)n + 1, sizeof(int64_t));
#pragma omp parallel
  {
    int64_t* mark = (int64_t*)malloc(sizeof(int64_t) * (size_t)(m ? m : 1));
    for (int64_t j = 0; j < m; ++j) mark[j] = -1;
#pragma omp for schedule(dynamic, 256)
    for (int64_t i = 0; i < n; ++i) {
      int64_t c = 0;
      for (int64_t p = A->rp[i]; p < A->rp[i + 1]; ++p) {
        int32_t k = A->ci[p];
        for (int64_t q = B->rp[k]; q < B->rp[k + 1]; ++q)
          if (mark[B->ci[q]] != i) {
            mark[B->ci[q]] = i;
            ++c;
          }
      }
      cnt[i + 1] = c;
    }
    free(mark);
  }
  for (int64_t i = 0; i < n; ++i) cnt[i + 1] += cnt[i];
  C->n = n;
  C->m = m;
  C->nnz = cnt[n];
  C->rp = cnt;
  C->ci = (int32_t*)malloc(sizeof(int32_t) * (size_t)(C->nnz ? C->nnz : 1));
  C->v = (float*)malloc(sizeof(float) * (size_t)(C->nnz ? C->nnz : 1));
#pragma omp parallel
  {
    float* acc = (float*)calloc((size_t)(m ? m : 1), sizeof(float));
    int64_t* mark = (int64_t*)malloc(sizeof(int64_t) * (size_t)(m ? m : 1));
    for (int64_t j = 0; j < m; ++j) mark[j] = -1;
#pragma omp for schedule(dynamic, 256)
    for (int64_t i = 0; i < n; ++i) {
      int64_t o = C->rp[i], c = 0;
      for (int64_t p = A->rp[i]; p < A->rp[i + 1]; ++p) {
        int32_t k = A->ci[p];
        float a = A->v[p];
        for (int64_t q = B->rp[k]; q < B->rp[k + 1]; ++q) {
          int32_t j = B->ci[q];
          if (mark[j] != i) {
            mark[j] = i;
            C->ci[o + c++] = j;
            acc[j] = 0.0f;
          }
          acc[j] += a * B->v[q];
        }
      }
      qsort(C->ci + o, (size_t)c, sizeof(int32_t), cmp_int32);
      for (int64_t t = 0; t < c; ++t) C->v[o + t] = acc[C->ci[o + t]];
    }
    free(acc);
    free(mark);
  }
}

static void fcsr_transpose(const fcsr* A, fcsr* T) {
  T->n = A->m;
  T->m = A->n;
  T->nnz = A->nnz;
  T->rp = (int64_t*)calloc((size_t)T->n + 1, sizeof(int64_t));
  T->ci = (int32_t*)malloc(sizeof(int32_t) * (size_t)(A->nnz ? A->nnz : 1));
  T->v = (float*)malloc(sizeof(float) * (size_t)(A->nnz ? A->nnz : 1));
  for (int64_t p = 0; p < A->nnz; ++p) T->rp[A->ci[p] + 1]++;
  for (int64_t i = 0; i < T->n; ++i) T->rp[i + 1] += T->rp[i];
  int64_t* pos = (int64_t*)malloc(sizeof(int64_t) * (size_t)(T->n ? T->n : 1));
  for (int64_t i = 0; i < T->n; ++i) pos[i] = T->rp[i];
  for (int64_t i = 0; i < A->n; ++i)
    for (int64_t p = A->rp[i]; p < A->rp[i + 1]; ++p) {
      int64_t d = pos[A->ci[p]]++;
      T->ci[d] = (int32_t)i;
      T->v[d] = A->v[p];
    }
  free(pos);
}

/* graph of a matrix: pattern without the diagonal (ascending) */
static void graph_of(const fcsr* A, int32_t** gp, int32_t** gj) {
  int32_t* p = (int32_t*)malloc(sizeof(int32_t) * (size_t)(A->n + 1));
  int32_t* j = (int32_t*)malloc(sizeof(int32_t) * (size_t)(A->nnz ? A->nnz : 1));
  p[0] = 0;
  int64_t c = 0;
  for (int64_t i = 0; i < A->n; ++i) {
    for (int64_t q = A->rp[i]; q < A->rp[i + 1]; ++q)
      if (A->ci[q] != (int32_t)i) j[c++] = A->ci[q];
    p[i + 1] = (int32_t)c;
  }
  *gp = p;
  *gj = j;
}

/* induced aggregate graph (misHelpers::getInducedGraph, misHelpers.cu:1078):
 * aggregates adjacent through any fine edge, ascending, without self */
void orc_induced_graph(int64_t n, const int32_t* gp, const int32_t* gj, const int32_t* agg, int64_t nc, int32_t** cp,
                       int32_t** cj) {
  int64_t* mptr = (int64_t*)calloc((size_t)nc + 1, sizeof(int64_t));
  for (int64_t i = 0; i < n; ++i) mptr[agg[i] + 1]++;
  for (int64_t a = 0; a < nc; ++a) mptr[a + 1] += mptr[a];
  int32_t* mem = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n ? n : 1));
  int64_t* pos = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nc ? nc : 1));
  for (int64_t a = 0; a < nc; ++a) pos[a] = mptr[a];
  for (int64_t i = 0; i < n; ++i) mem[pos[agg[i]]++] = (int32_t)i;
  int32_t* p = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nc + 1));
  avec* rows = (avec*)calloc((size_t)(nc ? nc : 1), sizeof(avec));
#pragma omp parallel
  {
    int64_t* mark = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nc ? nc : 1));
    for (int64_t a = 0; a < nc; ++a) mark[a] = -1;
#pragma omp for schedule(dynamic, 64)
    for (int64_t a = 0; a < nc; ++a) {
      for (int64_t t = mptr[a]; t < mptr[a + 1]; ++t) {
        int32_t i = mem[t];
        for (int32_t q = gp[i]; q < gp[i + 1]; ++q) {
          int32_t c = agg[gj[q]];
          if (c != (int32_t)a && mark[c] != a) {
            mark[c] = a;
            avec_push(&rows[a], c);
          }
        }
      }
      if (rows[a].n) qsort(rows[a].v, (size_t)rows[a].n, sizeof(int32_t), cmp_int32);
    }
    free(mark);
  }
  p[0] = 0;
  for (int64_t a = 0; a < nc; ++a) p[a + 1] = p[a] + rows[a].n;
  int32_t* j = (int32_t*)malloc(sizeof(int32_t) * (size_t)(p[nc] ? p[nc] : 1));
  for (int64_t a = 0; a < nc; ++a) {
    if (rows[a].n) memcpy(j + p[a], rows[a].v, sizeof(int32_t) * (size_t)rows[a].n);
    free(rows[a].v);
  }
  free(rows);
  free(mptr);
  free(mem);
  free(pos);
  *cp = p;
  *cj = j;
}

void orc_free(void* p) { free(p); }

/* aggregateWeightedGraph (misHelpers.cu:603-677): MIS-1 roots on the
 * aggregate graph, first-neighbour allocation, one removal of partitions
 * lighter than 0.7 x the average weight (removeRuntyPartitions :783-820),
 * re-allocation; partitions heavier than part_max are then split in aggregate
 * order (deviation, see the file header). */
static int64_t partition_aggregates(int64_t nc, const int32_t* cp, const int32_t* cj, const int32_t* w, int64_t full,
                                    int part_max, uint64_t seed, int32_t* part) {
  if (nc == 0) return 0;
  int64_t np = orc_aggregate(nc, cp, cj, 1, 1, 0, seed, part);
  int64_t* ws = (int64_t*)calloc((size_t)np, sizeof(int64_t));
  for (int64_t a = 0; a < nc; ++a) ws[part[a]] += w[a];
  int64_t thr = (int64_t)((double)full / (double)np * 0.7);
  int32_t* aggd = (int32_t*)malloc(sizeof(int32_t) * (size_t)nc);
  int32_t* ex = (int32_t*)malloc(sizeof(int32_t) * (size_t)np);
  int32_t removed = 0;
  for (int64_t p = 0; p < np; ++p) {
    ex[p] = removed;
    removed += ws[p] < thr;
  }
  if (removed && removed < np) {
    for (int64_t a = 0; a < nc; ++a) {
      int32_t p = part[a];
      if (ws[p] < thr) {
        part[a] = -1;
        aggd[a] = 0;
      } else {
        part[a] = p - ex[p];
        aggd[a] = 1;
      }
    }
    int32_t* pout = (int32_t*)malloc(sizeof(int32_t) * (size_t)nc);
    int64_t prev = -1;
    for (;;) {
      int64_t un = 0;
      for (int64_t a = 0; a < nc; ++a) {
        pout[a] = part[a];
        if (aggd[a]) continue;
        for (int32_t q = cp[a]; q < cp[a + 1]; ++q)
          if (part[cj[q]] != -1) {
            pout[a] = part[cj[q]];
            break;
          }
      }
      for (int64_t a = 0; a < nc; ++a) {
        part[a] = pout[a];
        if (part[a] != -1) aggd[a] = 1;
        else ++un;
      }
      if (un == 0) break;
      if (un == prev) {
        int32_t mx = -1;
        for (int64_t a = 0; a < nc; ++a)
          if (part[a] > mx) mx = part[a];
        for (int64_t a = 0; a < nc; ++a)
          if (part[a] == -1) part[a] = ++mx;
        break;
      }
      prev = un;
    }
    free(pout);
    np = 0;
    for (int64_t a = 0; a < nc; ++a)
      if (part[a] + 1 > np) np = part[a] + 1;
  }
  free(ws);
  free(aggd);
  free(ex);
  int64_t* cur_w = (int64_t*)calloc((size_t)np, sizeof(int64_t));
  int32_t* cur_id = (int32_t*)malloc(sizeof(int32_t) * (size_t)np);
  for (int64_t p = 0; p < np; ++p) cur_id[p] = -1;
  int32_t next = 0;
  for (int64_t a = 0; a < nc; ++a) {
    int32_t p = part[a];
    if (cur_id[p] < 0 || (cur_w[p] + w[a] > part_max && cur_w[p] > 0)) {
      cur_id[p] = next++;
      cur_w[p] = 0;
    }
    cur_w[p] += w[a];
    part[a] = cur_id[p];
  }
  free(cur_w);
  free(cur_id);
  return next;
}

static void lu_factor_f(float* a, int32_t* piv, int64_t n) {
  for (int64_t j = 0; j < n; ++j) {
    int64_t p = j;
    for (int64_t i = j + 1; i < n; ++i)
      if (fabsf(a[i * n + j]) > fabsf(a[p * n + j])) p = i;
    piv[j] = (int32_t)p;
    if (p != j)
      for (int64_t k = 0; k < n; ++k) {
        float t = a[j * n + k];
        a[j * n + k] = a[p * n + k];
        a[p * n + k] = t;
      }
    float d = a[j * n + j];
    if (d == 0.0f) continue;
    for (int64_t i = j + 1; i < n; ++i) {
      float f = a[i * n + j] / d;
      a[i * n + j] = f;
      for (int64_t k = j + 1; k < n; ++k) a[i * n + k] -= f * a[j * n + k];
    }
  }
}

static void lu_solve_f(const float* a, const int32_t* piv, int64_t n, const float* b, float* x) {
  for (int64_t i = 0; i < n; ++i) x[i] = b[i];
  for (int64_t j = 0; j < n; ++j)
    if (piv[j] != j) {
      float t = x[j];
      x[j] = x[piv[j]];
      x[piv[j]] = t;
    }
  for (int64_t i = 0; i < n; ++i) {
    float s = x[i];
    for (int64_t k = 0; k < i; ++k) s -= a[i * n + k] * x[k];
    x[i] = s;
  }
  for (int64_t i = n - 1; i >= 0; --i) {
    float s = x[i];
    for (int64_t k = i + 1; k < n; ++k) s -= a[i * n + k] * x[k];
    x[i] = s / a[i * n + i];
  }
}

typedef struct {
  olevel* L;
  int nlev;
  float* lu;
  int32_t* piv;
  int64_t nc;
  int pre, post, relaxes;
  float w;
} ohier;

/* nu Jacobi sweeps of x on the in-partition block for right-hand side bl */
static void block_sweeps(olevel* L, const float* bl, int nu, float w) {
  const fcsr* A = &L->A;
#pragma omp parallel for schedule(dynamic, 4)
  for (int64_t p = 0; p < L->nparts; ++p) {
    for (int it = 0; it < nu; ++it) {
      for (int64_t t = L->pptr[p]; t < L->pptr[p + 1]; ++t) {
        int32_t i = L->prow[t];
        float s = 0.0f;
        for (int64_t q = A->rp[i]; q < A->rp[i + 1]; ++q) {
          int32_t j = A->ci[q];
          if (j != i && L->part_of[j] == (int32_t)p) s += A->v[q] * L->x[j];
        }
        L->y[i] = s;
      }
      for (int64_t t = L->pptr[p]; t < L->pptr[p + 1]; ++t) {
        int32_t i = L->prow[t];
        L->x[i] += w * (bl[i] - L->y[i] - L->diag[i] * L->x[i]) / L->diag[i];
      }
    }
  }
}

static void ocycle(ohier* H, int l) {
  olevel* L = &H->L[l];
  const fcsr* A = &L->A;
  const int64_t n = A->n;
  if (l == H->nlev - 1) {
    lu_solve_f(H->lu, H->piv, H->nc, L->b, L->x);
    return;
  }
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) L->x[i] = H->w * L->b[i] / L->diag[i];
  block_sweeps(L, L->b, H->pre, H->w);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    float s = 0.0f;
    for (int64_t q = A->rp[i]; q < A->rp[i + 1]; ++q) s += A->v[q] * L->x[A->ci[q]];
    L->r[i] = L->b[i] - s;
  }
  olevel* N = &H->L[l + 1];
#pragma omp parallel for schedule(static)
  for (int64_t I = 0; I < L->R.n; ++I) {
    float s = 0.0f;
    for (int64_t q = L->R.rp[I]; q < L->R.rp[I + 1]; ++q) s += L->R.v[q] * L->r[L->R.ci[q]];
    N->b[I] = s;
  }
  ocycle(H, l + 1);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    float s = 0.0f;
    for (int64_t q = L->P.rp[i]; q < L->P.rp[i + 1]; ++q) s += L->P.v[q] * N->x[L->P.ci[q]];
    L->x[i] += s;
  }
  for (int rel = 0; rel < H->relaxes; ++rel) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
      float s = 0.0f;
      for (int64_t q = A->rp[i]; q < A->rp[i + 1]; ++q) {
        int32_t j = A->ci[q];
        if (L->part_of[j] != L->part_of[i]) s += A->v[q] * L->x[j];
      }
      L->xo[i] = L->b[i] - s;
    }
    block_sweeps(L, L->xo, H->post, H->w);
  }
}

static void ohier_free(ohier* H) {
  for (int l = 0; l < H->nlev; ++l) {
    olevel* L = &H->L[l];
    fcsr_free(&L->A);
    fcsr_free(&L->P);
    fcsr_free(&L->R);
    free(L->diag);
    free(L->agg);
    free(L->pptr);
    free(L->prow);
    free(L->part_of);
    free(L->x);
    free(L->b);
    free(L->r);
    free(L->y);
    free(L->xo);
  }
  free(H->L);
  free(H->lu);
  free(H->piv);
}

int orc_amg_pcg(int64_t n, const int64_t* rp, const int32_t* ci, const double* val, const double* b, double* x,
                double tol, int max_iters, int top_size, int max_levels, int pre, int post, int relaxes,
                double smoother_weight, double pro_omega, int part_max, uint64_t seed, double* out) {
  double t0 = now_s();
  ohier H;
  memset(&H, 0, sizeof(H));
  H.pre = pre;
  H.post = post;
  H.relaxes = relaxes;
  H.w = (float)smoother_weight;
  if (max_levels < 1) max_levels = 1;
  H.L = (olevel*)calloc((size_t)max_levels + 1, sizeof(olevel));
  olevel* L0 = &H.L[0];
  L0->A.n = L0->A.m = n;
  L0->A.nnz = rp[n];
  L0->A.rp = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n + 1));
  memcpy(L0->A.rp, rp, sizeof(int64_t) * (size_t)(n + 1));
  L0->A.ci = (int32_t*)malloc(sizeof(int32_t) * (size_t)(rp[n] ? rp[n] : 1));
  memcpy(L0->A.ci, ci, sizeof(int32_t) * (size_t)rp[n]);
  L0->A.v = (float*)malloc(sizeof(float) * (size_t)(rp[n] ? rp[n] : 1));
  for (int64_t q = 0; q < rp[n]; ++q) L0->A.v[q] = (float)val[q];
  int32_t *gp, *gj;
  graph_of(&L0->A, &gp, &gj);
  const int fine_depth = 2, min_size = 9;
  int64_t tot_nnz = 0;
  int l = 0;
  for (;; ++l) {
    olevel* L = &H.L[l];
    const int64_t nl = L->A.n;
    tot_nnz += L->A.nnz;
    L->diag = (float*)malloc(sizeof(float) * (size_t)nl);
    for (int64_t i = 0; i < nl; ++i) {
      L->diag[i] = 0.0f;
      for (int64_t q = L->A.rp[i]; q < L->A.rp[i + 1]; ++q)
        if (L->A.ci[q] == (int32_t)i) L->diag[i] = L->A.v[q];
    }
    L->x = (float*)calloc((size_t)nl, sizeof(float));
    L->b = (float*)calloc((size_t)nl, sizeof(float));
    L->r = (float*)calloc((size_t)nl, sizeof(float));
    L->y = (float*)calloc((size_t)nl, sizeof(float));
    L->xo = (float*)calloc((size_t)nl, sizeof(float));
    if (nl < top_size || l + 1 >= max_levels) break;
    L->agg = (int32_t*)malloc(sizeof(int32_t) * (size_t)nl);
    L->nc = orc_aggregate(nl, gp, gj, fine_depth, min_size, 0, seed + 0x9E37ull * (uint64_t)l, L->agg);
    if (L->nc >= nl || L->nc == 0) break;
    int32_t *cp, *cj;
    orc_induced_graph(nl, gp, gj, L->agg, L->nc, &cp, &cj);
    int32_t* wsz = (int32_t*)calloc((size_t)L->nc, sizeof(int32_t));
    for (int64_t i = 0; i < nl; ++i) wsz[L->agg[i]]++;
    int32_t* pa = (int32_t*)malloc(sizeof(int32_t) * (size_t)L->nc);
    if (part_max > 0) {
      L->nparts = partition_aggregates(L->nc, cp, cj, wsz, nl, part_max, seed ^ 0xC0A5Eull, pa);
    } else { /* diagnostic: one partition (global weighted Jacobi) */
      for (int64_t a = 0; a < L->nc; ++a) pa[a] = 0;
      L->nparts = 1;
    }
    L->part_of = (int32_t*)malloc(sizeof(int32_t) * (size_t)nl);
    L->pptr = (int64_t*)calloc((size_t)L->nparts + 1, sizeof(int64_t));
    L->prow = (int32_t*)malloc(sizeof(int32_t) * (size_t)nl);
    for (int64_t i = 0; i < nl; ++i) {
      L->part_of[i] = pa[L->agg[i]];
      L->pptr[L->part_of[i] + 1]++;
    }
    for (int64_t p = 0; p < L->nparts; ++p) L->pptr[p + 1] += L->pptr[p];
    {
      int64_t* pos = (int64_t*)malloc(sizeof(int64_t) * (size_t)L->nparts);
      for (int64_t p = 0; p < L->nparts; ++p) pos[p] = L->pptr[p];
      for (int64_t i = 0; i < nl; ++i) L->prow[pos[L->part_of[i]]++] = (int32_t)i;
      free(pos);
    }
    free(wsz);
    free(pa);
    /* P = (I - omega D^-1 A) T in float: S entries in ascending column order,
     * then the tentative 1 (order of the reference's reduce_by_key) */
    {
      fcsr* P = &L->P;
      P->n = nl;
      P->m = L->nc;
      P->rp = (int64_t*)calloc((size_t)nl + 1, sizeof(int64_t));
      int32_t* tc = (int32_t*)malloc(sizeof(int32_t) * (size_t)(L->A.nnz + nl));
      float* tv = (float*)malloc(sizeof(float) * (size_t)(L->A.nnz + nl));
      int64_t c = 0;
      const float lam = (float)pro_omega;
      for (int64_t i = 0; i < nl; ++i) {
        int64_t o = c;
        for (int64_t q = L->A.rp[i]; q < L->A.rp[i + 1]; ++q) {
          int32_t J = L->agg[L->A.ci[q]];
          float v = (-lam * L->A.v[q]) / L->diag[i];
          int64_t t = o;
          while (t < c && tc[t] != J) ++t;
          if (t == c) {
            tc[c] = J;
            tv[c] = v;
            ++c;
          } else {
            tv[t] += v;
          }
        }
        int64_t t = o;
        while (t < c && tc[t] != L->agg[i]) ++t;
        if (t == c) {
          tc[c] = L->agg[i];
          tv[c] = 1.0f;
          ++c;
        } else {
          tv[t] += 1.0f;
        }
        for (int64_t a = o + 1; a < c; ++a) {
          int32_t kc = tc[a];
          float kv = tv[a];
          int64_t bb = a - 1;
          while (bb >= o && tc[bb] > kc) {
            tc[bb + 1] = tc[bb];
            tv[bb + 1] = tv[bb];
            --bb;
          }
          tc[bb + 1] = kc;
          tv[bb + 1] = kv;
        }
        P->rp[i + 1] = c;
      }
      P->nnz = c;
      P->ci = tc;
      P->v = tv;
    }
    fcsr_transpose(&L->P, &L->R);
    fcsr AP;
    fcsr_spgemm(&L->A, &L->P, &AP);
    fcsr_spgemm(&L->R, &AP, &H.L[l + 1].A);
    fcsr_free(&AP);
    free(gp);
    free(gj);
    gp = cp;
    gj = cj;
  }
  free(gp);
  free(gj);
  H.nlev = l + 1;
  {
    olevel* C = &H.L[l];
    H.nc = C->A.n;
    const int64_t nn = H.nc * H.nc;
    H.lu = (float*)calloc((size_t)(nn > 0 ? nn : 1), sizeof(float));
    H.piv = (int32_t*)malloc(sizeof(int32_t) * (size_t)(H.nc ? H.nc : 1));
    for (int64_t i = 0; i < H.nc; ++i)
      for (int64_t q = C->A.rp[i]; q < C->A.rp[i + 1]; ++q) H.lu[i * H.nc + C->A.ci[q]] += C->A.v[q];
    lu_factor_f(H.lu, H.piv, H.nc);
  }
  double t1 = now_s();
  double* r = (double*)malloc(sizeof(double) * (size_t)n);
  double* z = (double*)malloc(sizeof(double) * (size_t)n);
  double* p = (double*)malloc(sizeof(double) * (size_t)n);
  double* q = (double*)malloc(sizeof(double) * (size_t)n);
  double bnorm = 0.0;
  for (int64_t i = 0; i < n; ++i) bnorm += b[i] * b[i];
  bnorm = sqrt(bnorm);
  int iters = 0, converged = 0, restarts = 0;
  double rel = 0.0;
  for (;;) {
    double rr = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : rr)
    for (int64_t i = 0; i < n; ++i) {
      double s = 0.0;
      for (int64_t t = rp[i]; t < rp[i + 1]; ++t) s += val[t] * x[ci[t]];
      r[i] = b[i] - s;
      rr += r[i] * r[i];
    }
    rel = bnorm > 0 ? sqrt(rr) / bnorm : 0.0;
    if (rel <= tol) {
      converged = 1;
      break;
    }
    if (iters >= max_iters || restarts > 3) break;
    ++restarts;
    for (int64_t i = 0; i < n; ++i) L0->b[i] = (float)r[i];
    ocycle(&H, 0);
    double rz = 0.0;
    for (int64_t i = 0; i < n; ++i) {
      z[i] = L0->x[i];
      p[i] = z[i];
      rz += r[i] * z[i];
    }
    while (iters < max_iters) {
      double pq = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : pq)
      for (int64_t i = 0; i < n; ++i) {
        double s = 0.0;
        for (int64_t t = rp[i]; t < rp[i + 1]; ++t) s += val[t] * p[ci[t]];
        q[i] = s;
        pq += s * p[i];
      }
      double alpha = rz / pq;
      double nr = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : nr)
      for (int64_t i = 0; i < n; ++i) {
        x[i] += alpha * p[i];
        r[i] -= alpha * q[i];
        nr += r[i] * r[i];
      }
      ++iters;
      if (sqrt(nr) / bnorm <= tol) break;
#pragma omp parallel for schedule(static)
      for (int64_t i = 0; i < n; ++i) L0->b[i] = (float)r[i];
      ocycle(&H, 0);
      double rzn = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : rzn)
      for (int64_t i = 0; i < n; ++i) {
        z[i] = L0->x[i];
        rzn += z[i] * r[i];
      }
      double beta = rzn / rz;
      rz = rzn;
#pragma omp parallel for schedule(static)
      for (int64_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
    }
  }
  double t2 = now_s();
  int threads = 1;
#ifdef _OPENMP
  threads = omp_get_max_threads();
#endif
  out[0] = iters;
  out[1] = t1 - t0;
  out[2] = t2 - t1;
  out[3] = rel;
  out[4] = threads;
  out[5] = H.nlev;
  out[6] = (double)tot_nnz / (double)(rp[n] ? rp[n] : 1);
  out[7] = converged;
  free(r);
  free(z);
  free(p);
  free(q);
  ohier_free(&H);
  return converged ? 0 : 5;
}
