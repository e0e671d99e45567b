/*
 * solvers.c -- oracle linear solvers.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * The reference solves with HYPRE 3.0.0 (not vendored, not importable here):
 * PCG + BoomerAMG (LP/src/hypre_backend.cu:3-114, GSV1 main.cu:2224-2273,
 * 2617-2676) and BiCGSTAB + DiagScale (GSV1 main.cu:2553-2576,
 * LS/src/scalar_transport_library.cu:359-471).  HYPRE's iteration histories
 * are unpinnable here; only the converged solution is.  This file defines the
 * replacement algorithm (the GPU library runs the same one):
 *
 *   PCG, two-norm stop ||r|| <= max(relTol*||b||, absTol), x0 = caller's x,
 *   b == 0 -> x = 0 (HYPRE convention), preconditioned by one V-cycle of an
 *   unsmoothed-aggregation AMG:
 *     - coupling s_ij = -a_ij, strong when s_ij > 0 and s_ij >= theta*max_k s_ik
 *     - pairwise handshake matching (6 synchronous rounds): every unmatched
 *       row proposes its strongest unmatched strong neighbour, ties broken by
 *       a symmetric edge hash; mutual proposals pair up; a row left unmatched
 *       joins the pair of its strongest matched strong neighbour, or stays
 *       single when it has none
 *     - aggregate ids numbered in index order of their representative (the
 *       lower row of a pair, or the single row) -- deterministic
 *     - 'pairPasses' matchings per level (Galerkin sum in between)
 *     - coarse matrix = aggregate sum P^T A P, summed per coarse row in member
 *       order then CSR order
 *     - damped Jacobi pre/post smoothing (omega), or (smoother = 1) degree-2
 *       Chebyshev smoothing in D^-1 A on [chebFraction lmax, lmax], lmax its
 *       Gershgorin bound, on the chebLevels finest levels; dense Cholesky
 *       inverse on the coarsest level
 *     - optionally (kcycle) the K-cycle coarse correction on levels 1..kcycle.
 *   BiCGSTAB with right Jacobi preconditioning, same stop rule.
 *
 * Pinned by tests/test_oracle.py and tests/test_oracle_ops.py: PCG+AMG and
 * BiCGSTAB solutions agree with a scipy sparse direct solve (spsolve) to
 * 1e-9 / 1e-10 relative (SPD Laplacian; non-symmetric transport matrix);
 * aggregates partition the rows into connected sets with a coarsening
 * ratio above 3; the V-cycle is a linear, symmetric, positive operator, and
 * is the exact inverse when the coarsest level is the matrix itself (both
 * blocked-sum widths of the dense solve).  The K-cycle step satisfies the
 * Galerkin conditions of span{c1, c2} and composes with the fine level as a
 * scipy restatement does (test_kcycle_*).  The Chebyshev smoother: the
 * textbook recurrence (Saad, Algorithm 12.1) reproduces the closed-form
 * error polynomial T2((theta-lam)/delta)/T2(theta/delta) on the 1-D
 * Laplacian's eigenvectors, and this file's two-level cycle equals that
 * recurrence composed with an exact Galerkin coarse solve to 1e-11
 * (test_chebyshev_*).  The aggregation itself defines
 * the algorithm (it replaces HYPRE BoomerAMG), so its exact aggregates have
 * no outside reference: the GPU is checked against this file bit for bit.
 */
#include "oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* Loops whose iterations write disjoint outputs from inputs the loop does not
 * modify run on all threads (static schedule); every reduction, scan and
 * scatter stays serial, so the results are the same bits on any thread count. */
#define OR_PFOR _Pragma("omp parallel for schedule(static)")

void or_spmv(int n, const int *ro, const int *ci, const double *va,
             const double *x, double *y) {
  OR_PFOR
  for (int i = 0; i < n; ++i) {
    double s = 0.0;
    for (int k = ro[i]; k < ro[i + 1]; ++k) s += va[k] * x[ci[k]];
    y[i] = s;
  }
}

static double dotv(int n, const double *a, const double *b) {
  double s = 0.0;
  for (int i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

/* ------------------------------------------------------------- AMG */
uint32_t or_edge_hash(uint32_t a, uint32_t b);
uint32_t or_edge_hash(uint32_t a, uint32_t b) {
  uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
  uint32_t x = lo * 0x9E3779B1u ^ (hi + 0x7F4A7C15u);
  x ^= x >> 16;
  x *= 0x85EBCA6Bu;
  x ^= x >> 13;
  x *= 0xC2B2AE35u;
  x ^= x >> 16;
  return x;
}

typedef struct {
  int n;
  int *ro, *ci;
  double *va;
  int *diag;
  double *omdinv;
  int cheb;     /* degree-2 Chebyshev smoother on this level */
  double cc[4]; /* 1/theta, cv, cw, cr of the Chebyshev steps */
  int nc;
  int *agg;  /* n   -> coarse row */
  int *moff; /* nc+1 */
  int *mem;  /* n, members of each coarse row in ascending order */
  double *coarseInv; /* dense inverse on the coarsest level (n*n) */
  int isCoarsest;
} or_level;

#define OR_MATCH_ROUNDS 6
#define OR_GAL_CAP 256

/* one pairwise matching pass; returns number of aggregates, agg filled */
static int pair_match(int n, const int *ro, const int *ci, const double *va,
                      double theta, int *agg) {
  int *match = (int *)malloc(sizeof(int) * (size_t)n);
  int *cand = (int *)malloc(sizeof(int) * (size_t)n);
  double *smax = (double *)malloc(sizeof(double) * (size_t)n);
  OR_PFOR
  for (int i = 0; i < n; ++i) {
    match[i] = -1;
    double m = 0.0;
    for (int k = ro[i]; k < ro[i + 1]; ++k) {
      if (ci[k] == i) continue;
      double s = -va[k];
      if (s > m) m = s;
    }
    smax[i] = m;
  }
  for (int r = 0; r < OR_MATCH_ROUNDS; ++r) {
    OR_PFOR
    for (int i = 0; i < n; ++i) {
      cand[i] = -1;
      if (match[i] >= 0) continue;
      double bs = 0.0;
      uint32_t bh = 0;
      int best = -1;
      double thr = theta * smax[i];
      for (int k = ro[i]; k < ro[i + 1]; ++k) {
        int j = ci[k];
        if (j == i || match[j] >= 0) continue;
        double s = -va[k];
        if (!(s > 0.0) || s < thr) continue;
        uint32_t h = or_edge_hash((uint32_t)i, (uint32_t)j);
        if (best < 0 || s > bs || (s == bs && h > bh)) {
          best = j;
          bs = s;
          bh = h;
        }
      }
      cand[i] = best;
    }
    OR_PFOR
    for (int i = 0; i < n; ++i) {
      int j = cand[i];
      if (j >= 0 && cand[j] == i) match[i] = j;
    }
  }
  /* an unmatched row joins the pair of its strongest matched strong
   * neighbour (ties by edge hash); otherwise it stays a singleton */
  OR_PFOR
  for (int i = 0; i < n; ++i) {
    cand[i] = -1;
    if (match[i] >= 0) continue;
    double bs = 0.0;
    uint32_t bh = 0;
    int best = -1;
    double thr = theta * smax[i];
    for (int k = ro[i]; k < ro[i + 1]; ++k) {
      int j = ci[k];
      if (j == i || match[j] < 0) continue;
      double s = -va[k];
      if (!(s > 0.0) || s < thr) continue;
      uint32_t h = or_edge_hash((uint32_t)i, (uint32_t)j);
      if (best < 0 || s > bs || (s == bs && h > bh)) {
        best = j;
        bs = s;
        bh = h;
      }
    }
    cand[i] = best;
  }
  int nc = 0;
  for (int i = 0; i < n; ++i) {
    int rep = (match[i] >= 0) ? (i < match[i]) : (cand[i] < 0);
    if (rep) agg[i] = nc++;
  }
  for (int i = 0; i < n; ++i)
    if (match[i] >= 0 && match[i] < i) agg[i] = agg[match[i]];
  for (int i = 0; i < n; ++i)
    if (match[i] < 0 && cand[i] >= 0) {
      int j = cand[i];
      int r = match[j] < j ? match[j] : j;
      agg[i] = agg[r];
    }
  free(match);
  free(cand);
  free(smax);
  return nc;
}

/* members of each aggregate in ascending fine order (counting sort) */
static void build_members(int n, int nc, const int *agg, int **moff_out,
                          int **mem_out) {
  int *moff = (int *)calloc((size_t)nc + 1, sizeof(int));
  for (int i = 0; i < n; ++i) moff[agg[i] + 1]++;
  for (int I = 0; I < nc; ++I) moff[I + 1] += moff[I];
  int *mem = (int *)malloc(sizeof(int) * ((size_t)n + 1));
  int *fill = (int *)malloc(sizeof(int) * ((size_t)nc + 1));
  for (int I = 0; I < nc; ++I) fill[I] = moff[I];
  for (int i = 0; i < n; ++i) mem[fill[agg[i]]++] = i;
  free(fill);
  *moff_out = moff;
  *mem_out = mem;
}

/* coarse = P^T A P for piecewise-constant P (aggregate sum); the coarse
 * rows are independent (per-thread scratch), the row offsets a serial scan */
static int gal_row(const int *ro, const int *ci, const double *va, const int *agg,
                   const int *moff, const int *mem, int I, int fill, int **kJp,
                   double **kVp, int *capp) {
  int cnt = 0;
  for (int m = moff[I]; m < moff[I + 1]; ++m) {
    int i = mem[m];
    for (int k = ro[i]; k < ro[i + 1]; ++k) {
      int J = agg[ci[k]];
      int t = 0;
      for (; t < cnt; ++t)
        if ((*kJp)[t] == J) break;
      if (t == cnt) {
        if (cnt == *capp) {
          *capp *= 2;
          *kJp = (int *)realloc(*kJp, sizeof(int) * (size_t)*capp);
          *kVp = (double *)realloc(*kVp, sizeof(double) * (size_t)*capp);
        }
        (*kJp)[cnt] = J;
        if (fill) (*kVp)[cnt] = va[k];
        ++cnt;
      } else if (fill) {
        (*kVp)[t] += va[k];
      }
    }
  }
  return cnt;
}

static void galerkin(const int *ro, const int *ci, const double *va,
                     int nc, const int *agg, const int *moff, const int *mem,
                     int **cro_out, int **cci_out, double **cva_out) {
  int *cro = (int *)malloc(sizeof(int) * ((size_t)nc + 1));
  /* pass 1: count */
  cro[0] = 0;
#pragma omp parallel
  {
    int cap = 64;
    int *kJ = (int *)malloc(sizeof(int) * (size_t)cap);
    double *kV = (double *)malloc(sizeof(double) * (size_t)cap);
#pragma omp for schedule(static)
    for (int I = 0; I < nc; ++I)
      cro[I + 1] = gal_row(ro, ci, va, agg, moff, mem, I, 0, &kJ, &kV, &cap);
    free(kJ);
    free(kV);
  }
  for (int I = 0; I < nc; ++I) cro[I + 1] += cro[I];
  int *cci = (int *)malloc(sizeof(int) * ((size_t)cro[nc] + 1));
  double *cva = (double *)malloc(sizeof(double) * ((size_t)cro[nc] + 1));
  /* pass 2: accumulate in member/CSR order, then sort by column */
#pragma omp parallel
  {
    int cap = 64;
    int *kJ = (int *)malloc(sizeof(int) * (size_t)cap);
    double *kV = (double *)malloc(sizeof(double) * (size_t)cap);
#pragma omp for schedule(static)
    for (int I = 0; I < nc; ++I) {
      const int cnt = gal_row(ro, ci, va, agg, moff, mem, I, 1, &kJ, &kV, &cap);
      /* insertion sort by column */
      for (int a = 1; a < cnt; ++a) {
        int j = kJ[a];
        double v = kV[a];
        int b = a - 1;
        while (b >= 0 && kJ[b] > j) {
          kJ[b + 1] = kJ[b];
          kV[b + 1] = kV[b];
          --b;
        }
        kJ[b + 1] = j;
        kV[b + 1] = v;
      }
      for (int t = 0; t < cnt; ++t) {
        cci[cro[I] + t] = kJ[t];
        cva[cro[I] + t] = kV[t];
      }
    }
    free(kJ);
    free(kV);
  }
  *cro_out = cro;
  *cci_out = cci;
  *cva_out = cva;
}

/* dense Cholesky inverse; pivots <= 1e-13*max diag are treated as zero
 * (pseudo-inverse on that direction) */
static double *dense_inverse(int n, const int *ro, const int *ci,
                             const double *va) {
  double *A = (double *)calloc((size_t)n * n, sizeof(double));
  for (int i = 0; i < n; ++i)
    for (int k = ro[i]; k < ro[i + 1]; ++k) A[(size_t)i * n + ci[k]] += va[k];
  double dmax = 0.0;
  for (int i = 0; i < n; ++i)
    if (fabs(A[(size_t)i * n + i]) > dmax) dmax = fabs(A[(size_t)i * n + i]);
  /* L stored in the lower triangle of A */
  for (int k = 0; k < n; ++k) {
    double d = A[(size_t)k * n + k];
    for (int j = 0; j < k; ++j) d -= A[(size_t)k * n + j] * A[(size_t)k * n + j];
    if (d <= 1e-13 * dmax) {
      for (int i = k; i < n; ++i) A[(size_t)i * n + k] = 0.0;
      continue;
    }
    double l = sqrt(d);
    A[(size_t)k * n + k] = l;
    for (int i = k + 1; i < n; ++i) {
      double s = A[(size_t)i * n + k];
      for (int j = 0; j < k; ++j) s -= A[(size_t)i * n + j] * A[(size_t)k * n + j];
      A[(size_t)i * n + k] = s / l;
    }
  }
  /* M = L^-1 by forward substitution, column by column (a zero pivot gives
     a zero row of M), then inverse = M^T M: entry (i, j) is the sum over
     k >= max(i, j) of M[k][i] M[k][j] in increasing k (symmetric bit for
     bit) */
  double *M = (double *)calloc((size_t)n * n, sizeof(double));
  double *y = (double *)malloc(sizeof(double) * (size_t)n);
  for (int col = 0; col < n; ++col) {
    for (int i = 0; i < n; ++i) {
      double s = (i == col) ? 1.0 : 0.0;
      for (int j = 0; j < i; ++j) s -= A[(size_t)i * n + j] * y[j];
      double l = A[(size_t)i * n + i];
      y[i] = (l != 0.0) ? s / l : 0.0;
    }
    for (int i = 0; i < n; ++i) M[(size_t)i * n + col] = y[i];
  }
  free(y);
  free(A);
  double *inv = (double *)calloc((size_t)n * n, sizeof(double));
  OR_PFOR
  for (int i = 0; i < n; ++i)
    for (int j = 0; j <= i; ++j) {
      double s = 0.0;
      for (int k = i; k < n; ++k) s += M[(size_t)k * n + i] * M[(size_t)k * n + j];
      inv[(size_t)i * n + j] = s;
      inv[(size_t)j * n + i] = s;
    }
  free(M);
  return inv;
}

typedef struct {
  int nLevels;
  or_level lv[40];
  double **xb, **bb, **rb; /* per-level work vectors */
  double **kc, **kv;        /* K-cycle levels: first correction, its product */
  int sweeps, kcycle;
} or_amg;

static void amg_free(or_amg *H) {
  for (int l = 0; l < H->nLevels; ++l) {
    or_level *L = &H->lv[l];
    if (l > 0) {
      free(L->ro);
      free(L->ci);
      free(L->va);
    }
    free(L->diag);
    free(L->omdinv);
    free(L->agg);
    free(L->moff);
    free(L->mem);
    free(L->coarseInv);
    free(H->xb[l]);
    free(H->bb[l]);
    free(H->rb[l]);
    free(H->kc[l]);
    free(H->kv[l]);
  }
  free(H->xb);
  free(H->bb);
  free(H->rb);
  free(H->kc);
  free(H->kv);
}

static void level_diag(or_level *L, double omega) {
  L->diag = (int *)malloc(sizeof(int) * (size_t)L->n);
  L->omdinv = (double *)malloc(sizeof(double) * (size_t)L->n);
  for (int i = 0; i < L->n; ++i) {
    L->diag[i] = -1;
    for (int k = L->ro[i]; k < L->ro[i + 1]; ++k)
      if (L->ci[k] == i) L->diag[i] = k;
    double d = L->diag[i] >= 0 ? L->va[L->diag[i]] : 0.0;
    L->omdinv[i] = (d != 0.0) ? omega * (1.0 / d) : 0.0;
  }
}

/* Degree-2 Chebyshev smoothing for A with D = diag(A) (Adams, Brezina, Hu &
 * Tuminaro 2003, "Parallel multigrid smoothing: polynomial versus
 * Gauss-Seidel", and Saad, "Iterative Methods", 12.3).  lmax: the Gershgorin
 * bound max_i sum_k |a_ik| / |a_ii| of D^-1 A (a maximum: the same number in
 * any order of evaluation), lmin = fraction lmax; theta = (lmax + lmin)/2,
 * delta = (lmax - lmin)/2, sigma = theta/delta, rho1 = 1/sigma,
 * rho2 = 1/(2 sigma - rho1).  Steps from x0:
 *   x1 = x0 + D^-1 (b - A x0) / theta
 *   x2 = (1 + rho2 rho1) x1 - rho2 rho1 x0 + (2 rho2/delta) D^-1 (b - A x1)
 * kept as omdinv = (1/theta) / a_ii (the Jacobi sweep with omega = 1/theta)
 * and cc = {1/theta, cv = 1 + rho2 rho1, cw = -(rho2 rho1),
 * cr = (2 rho2/delta) theta}, so that the second step is
 * (cv x1 + cw x0) + cr (omdinv (b - A x1)). */
static void cheb_coefficients(double gmax, double fraction, double cc[4]) {
  const double lmax = gmax > 0.0 ? gmax : 1.0;
  const double lmin = fraction * lmax;
  const double theta = 0.5 * (lmax + lmin), delta = 0.5 * (lmax - lmin);
  const double sigma = theta / delta;
  const double rho1 = 1.0 / sigma;
  const double rho2 = 1.0 / (2.0 * sigma - rho1);
  cc[0] = 1.0 / theta;
  cc[1] = 1.0 + rho2 * rho1;
  cc[2] = -(rho2 * rho1);
  cc[3] = (2.0 * rho2 / delta) * theta;
}

static double gershgorin_dinv_a(const or_level *L) {
  double g = 0.0;
  for (int i = 0; i < L->n; ++i) {
    double s = 0.0, d = 0.0;
    for (int k = L->ro[i]; k < L->ro[i + 1]; ++k) {
      s += fabs(L->va[k]);
      if (L->ci[k] == i) d = L->va[k];
    }
    if (d != 0.0) {
      const double q = s / fabs(d);
      if (q > g) g = q;
    }
  }
  return g;
}

static void level_cheb(or_level *L, double fraction) {
  cheb_coefficients(gershgorin_dinv_a(L), fraction, L->cc);
  L->cheb = 1;
  free(L->diag);
  free(L->omdinv);
  level_diag(L, L->cc[0]);
}

static int level_is_cheb(const or_amg_options *o, int l, const or_level *L) {
  return o->smoother == 1 && !L->isCoarsest && (o->chebLevels <= 0 || l < o->chebLevels);
}

static void amg_setup(or_amg *H, int n, const int *ro, const int *ci,
                      const double *va, const or_amg_options *o) {
  memset(H, 0, sizeof(*H));
  H->sweeps = o->sweeps > 0 ? o->sweeps : 1;
  H->kcycle = o->kcycle > 0 ? o->kcycle : 0;
  int maxL = o->maxLevels > 1 ? o->maxLevels : 2;
  if (maxL > 40) maxL = 40;
  or_level *L = &H->lv[0];
  L->n = n;
  L->ro = (int *)ro;
  L->ci = (int *)ci;
  L->va = (double *)va;
  int l = 0;
  for (;;) {
    L = &H->lv[l];
    level_diag(L, o->jacobiOmega);
    if (L->n <= o->coarseSize || l == maxL - 1) {
      L->isCoarsest = 1;
      break;
    }
    /* pairPasses matchings composed */
    int *agg = (int *)malloc(sizeof(int) * (size_t)L->n);
    for (int i = 0; i < L->n; ++i) agg[i] = i;
    int curN = L->n;
    int *wro = L->ro, *wci = L->ci;
    double *wva = L->va;
    int owned = 0, wide = 0;
    int passes = o->pairPasses > 0 ? o->pairPasses : 1;
    for (int p = 0; p < passes && !wide; ++p) {
      int *a1 = (int *)malloc(sizeof(int) * (size_t)curN);
      int nc1 = pair_match(curN, wro, wci, wva, o->strength, a1);
      int *mo, *me;
      build_members(curN, nc1, a1, &mo, &me);
      int *nro, *nci;
      double *nva;
      galerkin(wro, wci, wva, nc1, a1, mo, me, &nro, &nci, &nva);
      free(mo);
      free(me);
      /* a coarse row wider than OR_GAL_CAP entries stops the hierarchy at
         this level (the GPU aggregate-sum kernel's row capacity) */
      for (int I = 0; I < nc1 && !wide; ++I)
        if (nro[I + 1] - nro[I] > OR_GAL_CAP) wide = 1;
      OR_PFOR
      for (int i = 0; i < L->n; ++i) agg[i] = a1[agg[i]];
      free(a1);
      if (owned) {
        free(wro);
        free(wci);
        free(wva);
      }
      wro = nro;
      wci = nci;
      wva = nva;
      owned = 1;
      curN = nc1;
    }
    if (wide || curN >= L->n || (double)curN > 0.9 * (double)L->n) {
      free(agg);
      free(wro);
      free(wci);
      free(wva);
      L->isCoarsest = 1;
      break;
    }
    L->nc = curN;
    L->agg = agg;
    build_members(L->n, curN, agg, &L->moff, &L->mem);
    or_level *C = &H->lv[l + 1];
    C->n = curN;
    C->ro = wro;
    C->ci = wci;
    C->va = wva;
    ++l;
  }
  H->nLevels = l + 1;
  for (int k = 0; k < H->nLevels; ++k)
    if (level_is_cheb(o, k, &H->lv[k])) level_cheb(&H->lv[k], o->chebFraction);
  or_level *Lc = &H->lv[l];
  if (Lc->n <= 4096) Lc->coarseInv = dense_inverse(Lc->n, Lc->ro, Lc->ci, Lc->va);
  H->xb = (double **)calloc((size_t)H->nLevels, sizeof(double *));
  H->bb = (double **)calloc((size_t)H->nLevels, sizeof(double *));
  H->rb = (double **)calloc((size_t)H->nLevels, sizeof(double *));
  H->kc = (double **)calloc((size_t)H->nLevels, sizeof(double *));
  H->kv = (double **)calloc((size_t)H->nLevels, sizeof(double *));
  for (int k = 0; k < H->nLevels; ++k) {
    H->xb[k] = (double *)calloc((size_t)H->lv[k].n, sizeof(double));
    H->bb[k] = (double *)calloc((size_t)H->lv[k].n, sizeof(double));
    H->rb[k] = (double *)calloc((size_t)H->lv[k].n, sizeof(double));
    if (k >= 1 && k <= H->kcycle && !H->lv[k].isCoarsest) {
      H->kc[k] = (double *)calloc((size_t)H->lv[k].n, sizeof(double));
      H->kv[k] = (double *)calloc((size_t)H->lv[k].n, sizeof(double));
    }
  }
}

/* 1 when some level takes the K-cycle correction (outer PCG then flexible) */
static int amg_flexible(const or_amg *H) {
  for (int k = 1; k < H->nLevels; ++k)
    if (H->kc[k]) return 1;
  return 0;
}

/* one damped Jacobi sweep x += omega D^-1 (b - A x) (Jacobi: uses old x) */
static void jacobi_sweep(const or_level *L, const double *b, double *x,
                         double *tmp) {
  OR_PFOR
  for (int i = 0; i < L->n; ++i) {
    double s = 0.0;
    for (int k = L->ro[i]; k < L->ro[i + 1]; ++k) s += L->va[k] * x[L->ci[k]];
    tmp[i] = x[i] + L->omdinv[i] * (b[i] - s);
  }
  memcpy(x, tmp, sizeof(double) * (size_t)L->n);
}

static void vcycle(or_amg *H, int l, const double *b, double *x);

/* K-cycle coarse correction on level c (Notay & Vassilevski 2008, "Recursive
 * Krylov-based multigrid cycles", two flexible-CG steps with the next cycle
 * as preconditioner, no early exit):
 *   c1 = B b, v1 = A c1, rho1 = c1.v1, a1 = c1.b / rho1
 *   b <- b - a1 v1 (in place: the level's right-hand side buffer)
 *   c2 = B b, w = A c2, g = c2.v1, d = c2.w, e = c2.b
 *   beta = d - g^2/rho1, a2 = e / beta
 *   x = (a1 - a2 g / rho1) c1 + a2 c2
 * with a zero rho1 (resp. beta <= 0) dropping the terms it divides. */
static void kcorrect(or_amg *H, int c, double *b, double *x) {
  or_level *L = &H->lv[c];
  const int n = L->n;
  double *c1 = H->kc[c], *v1 = H->kv[c], *w = H->rb[c];
  vcycle(H, c, b, c1);
  or_spmv(n, L->ro, L->ci, L->va, c1, v1);
  const double rho1 = dotv(n, c1, v1), g1 = dotv(n, c1, b);
  const double a1 = rho1 > 0.0 ? g1 / rho1 : 0.0;
  OR_PFOR
  for (int i = 0; i < n; ++i) b[i] = b[i] - a1 * v1[i];
  vcycle(H, c, b, x);
  or_spmv(n, L->ro, L->ci, L->va, x, w);
  const double g = dotv(n, x, v1), d = dotv(n, x, w), e = dotv(n, x, b);
  const double gr = rho1 > 0.0 ? g / rho1 : 0.0;
  const double beta = d - g * gr;
  const double a2 = beta > 0.0 ? e / beta : 0.0;
  const double al1 = a1 - a2 * gr;
  OR_PFOR
  for (int i = 0; i < n; ++i) x[i] = al1 * c1[i] + a2 * x[i];
}

static void vcycle(or_amg *H, int l, const double *b, double *x) {
  or_level *L = &H->lv[l];
  double *r = H->rb[l];
  if (L->isCoarsest) {
    if (L->coarseInv) {
      /* x = Ainv b summed in column chunks of 16 (n <= 1024) or 64: each
         chunk's partial sum in column order, then the chunk partials in
         chunk order (the blocked order the GPU coarse solve uses) */
      const int ch = L->n <= 1024 ? 16 : 64;
      OR_PFOR
      for (int i = 0; i < L->n; ++i) {
        double s = 0.0;
        const double *row = L->coarseInv + (size_t)i * L->n;
        for (int c = 0; c < L->n; c += ch) {
          double t = 0.0;
          const int e = c + ch < L->n ? c + ch : L->n;
          for (int j = c; j < e; ++j) t += row[j] * b[j];
          s += t;
        }
        x[i] = s;
      }
    } else {
      memset(x, 0, sizeof(double) * (size_t)L->n);
      for (int s = 0; s < 20; ++s) jacobi_sweep(L, b, x, r);
    }
    return;
  }
  /* pre-smoothing from zero: first sweep is x = omega D^-1 b (Chebyshev:
     the first step, omega = 1/theta) */
  OR_PFOR
  for (int i = 0; i < L->n; ++i) x[i] = L->omdinv[i] * b[i];
  if (L->cheb) {
    /* second step from x0 = 0: x2 = cv x1 + cr omdinv (b - A x1) */
    OR_PFOR
    for (int i = 0; i < L->n; ++i) {
      double s = 0.0;
      for (int k = L->ro[i]; k < L->ro[i + 1]; ++k) s += L->va[k] * x[L->ci[k]];
      const double t = L->omdinv[i] * (b[i] - s);
      r[i] = L->cc[1] * x[i] + L->cc[3] * t;
    }
    memcpy(x, r, sizeof(double) * (size_t)L->n);
  } else {
    for (int s = 1; s < H->sweeps; ++s) jacobi_sweep(L, b, x, r);
  }
  /* residual and restriction (member order) */
  OR_PFOR
  for (int i = 0; i < L->n; ++i) {
    double s = 0.0;
    for (int k = L->ro[i]; k < L->ro[i + 1]; ++k) s += L->va[k] * x[L->ci[k]];
    r[i] = b[i] - s;
  }
  double *bc = H->bb[l + 1];
  OR_PFOR
  for (int I = 0; I < L->nc; ++I) {
    double s = 0.0;
    for (int m = L->moff[I]; m < L->moff[I + 1]; ++m) s += r[L->mem[m]];
    bc[I] = s;
  }
  double *xc = H->xb[l + 1];
  if (H->kc[l + 1])
    kcorrect(H, l + 1, bc, xc);
  else
    vcycle(H, l + 1, bc, xc);
  OR_PFOR
  for (int i = 0; i < L->n; ++i) x[i] = x[i] + xc[L->agg[i]];
  if (L->cheb) {
    /* x0 = e (in r), x1 = e + omdinv (b - A e), x2 from x1 and e */
    memcpy(r, x, sizeof(double) * (size_t)L->n);
    double *x1 = (double *)malloc(sizeof(double) * (size_t)L->n);
    OR_PFOR
    for (int i = 0; i < L->n; ++i) {
      double s = 0.0;
      for (int k = L->ro[i]; k < L->ro[i + 1]; ++k) s += L->va[k] * r[L->ci[k]];
      x1[i] = r[i] + L->omdinv[i] * (b[i] - s);
    }
    OR_PFOR
    for (int i = 0; i < L->n; ++i) {
      double s = 0.0;
      for (int k = L->ro[i]; k < L->ro[i + 1]; ++k) s += L->va[k] * x1[L->ci[k]];
      const double t = L->omdinv[i] * (b[i] - s);
      x[i] = (L->cc[1] * x1[i] + L->cc[2] * r[i]) + L->cc[3] * t;
    }
    free(x1);
  } else {
    for (int s = 0; s < H->sweeps; ++s) jacobi_sweep(L, b, x, r);
  }
}

int or_amg_aggregate(int n, const int *ro, const int *ci, const double *va,
                     const or_amg_options *opt, int *agg) {
  or_amg H;
  or_amg_options o = *opt;
  o.maxLevels = 2;
  o.coarseSize = 0;
  amg_setup(&H, n, ro, ci, va, &o);
  int nc = 0;
  if (H.nLevels > 1) {
    memcpy(agg, H.lv[0].agg, sizeof(int) * (size_t)n);
    nc = H.lv[0].nc;
  } else {
    for (int i = 0; i < n; ++i) agg[i] = i;
    nc = n;
  }
  amg_free(&H);
  return nc;
}

int or_pcg_amg(int n, const int *ro, const int *ci, const double *va,
               const double *b, double *x, const or_amg_options *opt,
               double *finalRelRes, double *hist, int histCap) {
  or_amg H;
  amg_setup(&H, n, ro, ci, va, opt);
  double *r = (double *)malloc(sizeof(double) * (size_t)n);
  double *z = (double *)malloc(sizeof(double) * (size_t)n);
  double *p = (double *)malloc(sizeof(double) * (size_t)n);
  double *q = (double *)malloc(sizeof(double) * (size_t)n);
  double bnorm = sqrt(dotv(n, b, b));
  int it = 0, conv = 0;
  const int flex = amg_flexible(&H);
  double rnorm = 0.0, tol = 0.0, rho = 0.0;
  if (bnorm == 0.0) {
    memset(x, 0, sizeof(double) * (size_t)n);
    if (finalRelRes) *finalRelRes = 0.0;
    if (hist && histCap > 0) hist[0] = 0.0;
    conv = 1;
    goto done;
  }
  or_spmv(n, ro, ci, va, x, q);
  OR_PFOR
  for (int i = 0; i < n; ++i) r[i] = b[i] - q[i];
  rnorm = sqrt(dotv(n, r, r));
  tol = fmax(opt->relTol * bnorm, opt->absTol);
  if (hist && histCap > 0) hist[0] = rnorm;
  if (rnorm <= tol) {
    conv = 1;
    if (finalRelRes) *finalRelRes = rnorm / bnorm;
    goto done;
  }
  vcycle(&H, 0, r, z);
  memcpy(p, z, sizeof(double) * (size_t)n);
  rho = dotv(n, r, z);
  for (it = 1; it <= opt->maxIter; ++it) {
    or_spmv(n, ro, ci, va, p, q);
    double pq = dotv(n, p, q);
    if (!(pq > 0.0)) break;
    double alpha = rho / pq;
    OR_PFOR
    for (int i = 0; i < n; ++i) {
      x[i] += alpha * p[i];
      r[i] -= alpha * q[i];
    }
    rnorm = sqrt(dotv(n, r, r));
    if (hist && it < histCap) hist[it] = rnorm;
    if (rnorm <= tol) {
      conv = 1;
      break;
    }
    vcycle(&H, 0, r, z);
    double rhoNew = dotv(n, r, z);
    /* flexible (Polak-Ribiere) beta for the K-cycle, = rhoNew/rho for a
       fixed preconditioner: -(z.(r_new - r_old))/rho with r_new - r_old =
       -alpha q, alpha = rho/pq */
    double beta = flex ? -dotv(n, z, q) / pq : rhoNew / rho;
    OR_PFOR
    for (int i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
    rho = rhoNew;
  }
  if (it > opt->maxIter) it = opt->maxIter;
  if (finalRelRes) *finalRelRes = rnorm / bnorm;
done:
  free(r);
  free(z);
  free(p);
  free(q);
  amg_free(&H);
  return conv ? it : -it - 1;
}

/* ------------------------------------------------------------ BiCGSTAB */
int or_bicgstab_jacobi(int n, const int *ro, const int *ci, const double *va,
                       const double *b, double *x, int maxIter, double relTol,
                       double absTol, double *finalRelRes) {
  double *dinv = (double *)malloc(sizeof(double) * (size_t)n);
  OR_PFOR
  for (int i = 0; i < n; ++i) {
    double d = 0.0;
    for (int k = ro[i]; k < ro[i + 1]; ++k)
      if (ci[k] == i) d = va[k];
    dinv[i] = (d != 0.0) ? 1.0 / d : 1.0;
  }
  double *r = (double *)malloc(sizeof(double) * (size_t)n);
  double *r0 = (double *)malloc(sizeof(double) * (size_t)n);
  double *p = (double *)calloc((size_t)n, sizeof(double));
  double *v = (double *)calloc((size_t)n, sizeof(double));
  double *ph = (double *)malloc(sizeof(double) * (size_t)n);
  double *s = (double *)malloc(sizeof(double) * (size_t)n);
  double *sh = (double *)malloc(sizeof(double) * (size_t)n);
  double *t = (double *)malloc(sizeof(double) * (size_t)n);
  double bnorm = sqrt(dotv(n, b, b));
  int it = 0, conv = 0;
  double rnorm = 0.0, tol = 0.0;
  double rho = 1.0, alpha = 1.0, omega = 1.0;
  if (bnorm == 0.0) {
    memset(x, 0, sizeof(double) * (size_t)n);
    if (finalRelRes) *finalRelRes = 0.0;
    conv = 1;
    goto done;
  }
  or_spmv(n, ro, ci, va, x, t);
  OR_PFOR
  for (int i = 0; i < n; ++i) {
    r[i] = b[i] - t[i];
    r0[i] = r[i];
  }
  rnorm = sqrt(dotv(n, r, r));
  tol = fmax(relTol * bnorm, absTol);
  if (rnorm <= tol) {
    conv = 1;
    goto fin;
  }
  for (it = 1; it <= maxIter; ++it) {
    double rhoNew = dotv(n, r0, r);
    if (rhoNew == 0.0) break;
    double beta = (rhoNew / rho) * (alpha / omega);
    OR_PFOR
    for (int i = 0; i < n; ++i) p[i] = r[i] + beta * (p[i] - omega * v[i]);
    OR_PFOR
    for (int i = 0; i < n; ++i) ph[i] = dinv[i] * p[i];
    or_spmv(n, ro, ci, va, ph, v);
    double r0v = dotv(n, r0, v);
    if (r0v == 0.0) break;
    alpha = rhoNew / r0v;
    OR_PFOR
    for (int i = 0; i < n; ++i) s[i] = r[i] - alpha * v[i];
    double snorm = sqrt(dotv(n, s, s));
    if (snorm <= tol) {
      OR_PFOR
      for (int i = 0; i < n; ++i) x[i] += alpha * ph[i];
      rnorm = snorm;
      conv = 1;
      break;
    }
    OR_PFOR
    for (int i = 0; i < n; ++i) sh[i] = dinv[i] * s[i];
    or_spmv(n, ro, ci, va, sh, t);
    double tt = dotv(n, t, t);
    omega = (tt > 0.0) ? dotv(n, t, s) / tt : 0.0;
    OR_PFOR
    for (int i = 0; i < n; ++i) {
      x[i] += alpha * ph[i] + omega * sh[i];
      r[i] = s[i] - omega * t[i];
    }
    rnorm = sqrt(dotv(n, r, r));
    if (rnorm <= tol) {
      conv = 1;
      break;
    }
    if (omega == 0.0) break;
    rho = rhoNew;
  }
  if (it > maxIter) it = maxIter;
fin:
  if (finalRelRes) *finalRelRes = rnorm / bnorm;
done:
  free(dinv);
  free(r);
  free(r0);
  free(p);
  free(v);
  free(ph);
  free(s);
  free(sh);
  free(t);
  return conv ? it : -it - 1;
}

/* One V-cycle application z = M^-1 r with a freshly built hierarchy (test
 * hook; returns the number of levels). */
int or_amg_vcycle(int n, const int *ro, const int *ci, const double *va,
                  const or_amg_options *opt, const double *r, double *z) {
  or_amg H;
  amg_setup(&H, n, ro, ci, va, opt);
  vcycle(&H, 0, r, z);
  int nl = H.nLevels;
  amg_free(&H);
  return nl;
}

/* The K-cycle's two flexible-CG steps applied on the given matrix itself,
 * with the plain cycle of its own hierarchy as the inner preconditioner
 * (test hook for the Galerkin conditions of the step). */
int or_amg_kstep(int n, const int *ro, const int *ci, const double *va,
                 const or_amg_options *opt, const double *b, double *x) {
  or_amg H;
  or_amg_options o = *opt;
  o.kcycle = 0;
  amg_setup(&H, n, ro, ci, va, &o);
  H.kc[0] = (double *)calloc((size_t)n, sizeof(double));
  H.kv[0] = (double *)calloc((size_t)n, sizeof(double));
  double *bw = (double *)malloc(sizeof(double) * (size_t)n);
  memcpy(bw, b, sizeof(double) * (size_t)n);
  kcorrect(&H, 0, bw, x);
  free(bw);
  int nl = H.nLevels;
  amg_free(&H);
  return nl;
}

/* ---------------------------------------------- reusable PCG + AMG ---- */
/* Hierarchy kept between solves, as the reference keeps its BoomerAMG setup
 * between pAmgRebuildEvery rebuilds (GSV1 main.cu:3407-3411); new fine
 * values refresh the level-0 smoother only. */
typedef struct {
  or_amg H;
  or_amg_options opt;
  int n;
  const int *ro, *ci;
  const double *va;
  int built;
} or_pcg_handle;

void *or_pcg_create(int n, const int *ro, const int *ci,
                    const or_amg_options *opt) {
  or_pcg_handle *h = (or_pcg_handle *)calloc(1, sizeof(or_pcg_handle));
  h->n = n;
  h->ro = ro;
  h->ci = ci;
  h->opt = *opt;
  return h;
}

void or_pcg_set_values(void *hv, const double *va, int rebuild) {
  or_pcg_handle *h = (or_pcg_handle *)hv;
  h->va = va;
  if (rebuild || !h->built) {
    if (h->built) amg_free(&h->H);
    amg_setup(&h->H, h->n, h->ro, h->ci, va, &h->opt);
    h->built = 1;
  } else {
    or_level *L = &h->H.lv[0];
    L->va = (double *)va;
    free(L->diag);
    free(L->omdinv);
    level_diag(L, h->opt.jacobiOmega);
    if (L->cheb) level_cheb(L, h->opt.chebFraction);
  }
}

int or_pcg_solve(void *hv, const double *b, double *x, double relTol,
                 double absTol, int maxIter, double *finalRelRes) {
  or_pcg_handle *h = (or_pcg_handle *)hv;
  int n = h->n;
  const int *ro = h->ro, *ci = h->ci;
  const double *va = h->va;
  double *r = (double *)calloc((size_t)n, sizeof(double));
  double *z = (double *)malloc(sizeof(double) * (size_t)n);
  double *p = (double *)malloc(sizeof(double) * (size_t)n);
  double *q = (double *)malloc(sizeof(double) * (size_t)n);
  double bnorm = sqrt(dotv(n, b, b));
  int it = 0, conv = 0;
  const int flex = amg_flexible(&h->H);
  double rnorm = 0.0, tol = 0.0, rho = 0.0;
  if (bnorm == 0.0) {
    memset(x, 0, sizeof(double) * (size_t)n);
    if (finalRelRes) *finalRelRes = 0.0;
    conv = 1;
    goto done;
  }
  or_spmv(n, ro, ci, va, x, q);
  OR_PFOR
  for (int i = 0; i < n; ++i) r[i] = b[i] - q[i];
  rnorm = sqrt(dotv(n, r, r));
  tol = fmax(relTol * bnorm, absTol);
  if (rnorm <= tol) {
    conv = 1;
    if (finalRelRes) *finalRelRes = rnorm / bnorm;
    goto done;
  }
  vcycle(&h->H, 0, r, z);
  memcpy(p, z, sizeof(double) * (size_t)n);
  rho = dotv(n, r, z);
  for (it = 1; it <= maxIter; ++it) {
    or_spmv(n, ro, ci, va, p, q);
    double pq = dotv(n, p, q);
    if (!(pq > 0.0)) break;
    double alpha = rho / pq;
    OR_PFOR
    for (int i = 0; i < n; ++i) {
      x[i] += alpha * p[i];
      r[i] -= alpha * q[i];
    }
    rnorm = sqrt(dotv(n, r, r));
    if (rnorm <= tol) {
      conv = 1;
      break;
    }
    vcycle(&h->H, 0, r, z);
    double rhoNew = dotv(n, r, z);
    double beta = flex ? -dotv(n, z, q) / pq : rhoNew / rho;
    OR_PFOR
    for (int i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
    rho = rhoNew;
  }
  if (it > maxIter) it = maxIter;
  if (finalRelRes) *finalRelRes = rnorm / bnorm;
done:
  free(r);
  free(z);
  free(p);
  free(q);
  return conv ? it : -it - 1;
}

void or_pcg_destroy(void *hv) {
  or_pcg_handle *h = (or_pcg_handle *)hv;
  if (h->built) amg_free(&h->H);
  free(h);
}
