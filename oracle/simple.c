/*
 * simple.c -- oracle restatement of the generic_simple_v1 outer iteration.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Follows apps/generic_simple_v1/src/main.cu:3350-3634 (driver order) with
 * the operator restatements of assembly.c / simple_ops.c / gradient.c and
 * the replacement solvers of solvers.c (BiCGSTAB+Jacobi for momentum,
 * PCG + aggregation AMG with the pAmgRebuildEvery cadence for pressure).
 * Fields are host arrays; relchg/massRes are serial sums/maxima.
 *
 * Pinned by tests/test_oracle_ops.py (test_simple_*): zero boundary data
 * and zero fields stay exactly zero; a uniform-flow channel with slip-like
 * side walls stays uniform with p = 0 to round-off on a jittered mesh; on an
 * orthogonal channel the corrected flux satisfies discrete continuity to the
 * pressure-solve tolerance after every iteration, outflow balances inflow,
 * and the outer iteration converges.  And by tests/test_oracle_pins.py: on a
 * plane-Poiseuille channel the converged state equals the closed-form
 * solution of the discrete fully developed problem (velocity profile, zero
 * cross flow, p = G (x - L)) to 1e-8, for two different pairs of relaxation
 * factors, and that discrete solution tends to the analytic profile and
 * dp/dx = -8 mu U / H^2 at second order; three outer iterations on a
 * jittered channel (two velocity passes, three pressure passes) equal an
 * independent Python restatement of the driver order that composes the
 * pinned operators with scipy direct solves, to 1e-9.  Planted mistakes in
 * the driver (no pressure relaxation, a mis-scaled velocity correction) fail
 * these tests.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#include <omp.h>

typedef struct {
  double rho, mu;
  int nVelNonOrthCorr, nNonOrthCorr;
  double uRelax, pRelax, corrTol, tolMass, tolVel;
  int velMaxit;
  double velTol, velRelTol;
  int pMaxit;
  double pTol, pRelTol;
  int pAmgRebuildEvery;
  or_amg_options amg;
} or_simple_params;

typedef struct {
  double massRes, duRel, dvRel, dwRel, dpRel;
  int its[3], pcgLast, pcgTotal;
} or_simple_stats;

int or_simple_run(int nCells, int nFaces, int nIF, const int *owner,
                  const int *neigh, const int *bPatch, const int *ro,
                  const int *ci, const int *dp, const int *PP, const int *PN,
                  const int *NP, const int *NN, const double *cc,
                  const double *vol, const double *xf, const double *nf,
                  const double *Sf, const double *Af,
                  const or_simple_params *par, const int *patchUType,
                  const int *patchPType, const double *uBC,
                  const double *vBC, const double *wBC, const double *pBC,
                  int nIter, double *u, double *v, double *w, double *p,
                  double *phiOut, double *rAUout, or_simple_stats *stats);

static double relchg(int n, const double *a, const double *b) {
  double num = 0.0, den = 1.0; /* main.cu:3597-3601 */
  for (int i = 0; i < n; ++i) {
    double d = a[i] - b[i];
    num += d * d;
    den += a[i] * a[i];
  }
  return sqrt(num / den);
}

int or_simple_run(int nCells, int nFaces, int nIF, const int *owner,
                  const int *neigh, const int *bPatch, const int *ro,
                  const int *ci, const int *dp, const int *PP, const int *PN,
                  const int *NP, const int *NN, const double *cc,
                  const double *vol, const double *xf, const double *nf,
                  const double *Sf, const double *Af,
                  const or_simple_params *par, const int *patchUType,
                  const int *patchPType, const double *uBC,
                  const double *vBC, const double *wBC, const double *pBC,
                  int nIter, double *u, double *v, double *w, double *p,
                  double *phiOut, double *rAUout, or_simple_stats *stats) {
  const int nC = nCells, nF = nFaces, nBF = nFaces - nIF;
  const int64_t nnz = ro[nC];
  /* per-boundary-face BC views for the gradient */
  int *tU = (int *)malloc(sizeof(int) * (size_t)(nBF + 1));
  int *tP = (int *)malloc(sizeof(int) * (size_t)(nBF + 1));
  double *bv[3], *bp = (double *)malloc(sizeof(double) * (size_t)(nBF + 1));
  const double *fbc[3] = {uBC, vBC, wBC};
  for (int k = 0; k < 3; ++k) bv[k] = (double *)malloc(sizeof(double) * (size_t)(nBF + 1));
  for (int b = 0; b < nBF; ++b) {
    int pt = bPatch[nIF + b] - 1;
    tU[b] = patchUType[pt] == 1;
    tP[b] = patchPType[pt] == 1;
    for (int k = 0; k < 3; ++k) bv[k][b] = fbc[k][nIF + b];
    bp[b] = pBC[nIF + b];
  }
  double *q[3] = {u, v, w};
  double *qOld[3], *gp[3], *g = (double *)malloc(sizeof(double) * 3 * (size_t)nC);
  for (int k = 0; k < 3; ++k) {
    qOld[k] = (double *)malloc(sizeof(double) * (size_t)nC);
    gp[k] = (double *)malloc(sizeof(double) * (size_t)nC);
  }
  double *gpi = (double *)malloc(sizeof(double) * 3 * (size_t)nC);
  double *pOld = (double *)malloc(sizeof(double) * (size_t)nC);
  double *pCorr = (double *)calloc((size_t)nC, sizeof(double));
  double *rAU = (double *)calloc((size_t)nC, sizeof(double));
  double *qNew = (double *)malloc(sizeof(double) * (size_t)nC);
  double *vals = (double *)calloc((size_t)nnz, sizeof(double));
  double *pv = (double *)malloc(sizeof(double) * (size_t)nnz);
  double *rhs = (double *)malloc(sizeof(double) * (size_t)nC);
  double *phiStar = (double *)malloc(sizeof(double) * (size_t)nF);
  double *phi = (double *)malloc(sizeof(double) * (size_t)nF);
  double *phiNo = (double *)malloc(sizeof(double) * (size_t)nF);
  double *divStar = (double *)malloc(sizeof(double) * (size_t)nC);
  double *divCorr = (double *)calloc((size_t)nC, sizeof(double));
  double *divNo = (double *)malloc(sizeof(double) * (size_t)nC);
  void *ps = or_pcg_create(nC, ro, ci, &par->amg);
  int done = 0, diverged = 0;
  for (int step = 1; step <= nIter; ++step) {
    or_simple_stats st;
    memset(&st, 0, sizeof(st));
    for (int k = 0; k < 3; ++k) memcpy(qOld[k], q[k], sizeof(double) * (size_t)nC);
    memcpy(pOld, p, sizeof(double) * (size_t)nC);
    memset(pCorr, 0, sizeof(double) * (size_t)nC);
    /* grad(pOld) [bcP] */
    or_lsq_gradient_gsv1(nC, nF, nIF, owner, ro, ci, cc, xf, tP, bp, pOld, gpi);
    for (int i = 0; i < nC; ++i)
      for (int k = 0; k < 3; ++k) gp[k][i] = gpi[3 * i + k];
    /* momentum components (main.cu:3350-3384, 3401-3405) */
    const int nVel = (par->nVelNonOrthCorr > 0 ? par->nVelNonOrthCorr : 0) + 1;
    for (int comp = 0; comp < 3; ++comp) {
      memcpy(q[comp], qOld[comp], sizeof(double) * (size_t)nC);
      for (int it = 1; it <= nVel; ++it) {
        or_lsq_gradient_gsv1(nC, nF, nIF, owner, ro, ci, cc, xf, tU, bv[comp], q[comp], g);
        if (comp == 0 && it == 1) {
          or_momentum_assemble(nC, nF, nIF, owner, neigh, bPatch, dp, PP, PN, NP, NN, cc, vol,
                               xf, nf, Sf, Af, g, qOld[0], qOld[1], qOld[2], gp[comp],
                               patchUType, fbc[comp], patchUType, uBC, patchUType, vBC,
                               patchUType, wBC, par->rho, par->mu, 1.0, nnz, vals, rhs);
          or_momentum_relax(nC, dp, vals, rhs, q[comp], par->uRelax);
          or_extract_rau(nC, dp, vals, vol, rAU);
        } else {
          or_momentum_rhs_only(nC, nF, nIF, owner, neigh, bPatch, cc, vol, xf, nf, Sf, Af, g,
                               qOld[0], qOld[1], qOld[2], gp[comp], patchUType, fbc[comp],
                               patchUType, uBC, patchUType, vBC, patchUType, wBC, par->rho,
                               par->mu, 1.0, rhs);
          or_momentum_relax_rhs_only(nC, dp, vals, rhs, q[comp], par->uRelax);
        }
        memcpy(qNew, q[comp], sizeof(double) * (size_t)nC);
        double rr;
        int its = or_bicgstab_jacobi(nC, ro, ci, vals, rhs, qNew, par->velMaxit,
                                     par->velRelTol, par->velTol, &rr);
        st.its[comp] = its < 0 ? -its - 1 : its;
        double rc = relchg(nC, qNew, q[comp]);
        memcpy(q[comp], qNew, sizeof(double) * (size_t)nC);
        if (rc < par->corrTol) break;
      }
    }
    /* pressure matrix, AMG cadence (main.cu:3407-3411) */
    or_pressure_assemble(nC, nF, nIF, owner, neigh, bPatch, dp, PP, PN, NP, NN, cc, xf, nf, Af,
                         rAU, patchPType, nnz, pv);
    int every = par->pAmgRebuildEvery > 1 ? par->pAmgRebuildEvery : 1;
    or_pcg_set_values(ps, pv, step == 1 || ((step - 1) % every) == 0);
    /* predictor flux and its divergence (main.cu:3414-3445) */
    or_rhiechow_flux(nF, nIF, owner, neigh, bPatch, cc, xf, nf, Af, q[0], q[1], q[2], p, gpi,
                     rAU, patchUType, uBC, patchUType, vBC, patchUType, wBC, par->rho, phiStar);
    memcpy(phi, phiStar, sizeof(double) * (size_t)nF);
    or_continuity(nC, nF, nIF, owner, neigh, phi, divStar);
    const int nPass = (par->nNonOrthCorr > 0 ? par->nNonOrthCorr : 0) + 1;
    for (int pass = 1; pass <= nPass; ++pass) {
      or_lsq_gradient_gsv1(nC, nF, nIF, owner, ro, ci, cc, xf, tP, bp, pCorr, g);
      or_pressure_nonorth(nC, nF, nIF, owner, neigh, bPatch, cc, xf, nf, Sf, Af, rAU, patchPType,
                          g, phiNo, divNo);
      for (int c = 0; c < nC; ++c) rhs[c] = -divStar[c] + divNo[c];
      double rr;
      double relTol = (pass == nPass) ? 0.0 : par->pRelTol;
      int its = or_pcg_solve(ps, rhs, pCorr, relTol, par->pTol, par->pMaxit, &rr);
      its = its < 0 ? -its - 1 : its;
      st.pcgLast = its;
      st.pcgTotal += its;
      if (pass == nPass) {
        or_lsq_gradient_gsv1(nC, nF, nIF, owner, ro, ci, cc, xf, tP, bp, pCorr, g);
        or_pressure_nonorth(nC, nF, nIF, owner, neigh, bPatch, cc, xf, nf, Sf, Af, rAU,
                            patchPType, g, phiNo, divNo);
        or_correct_flux(nF, nIF, owner, neigh, bPatch, cc, xf, nf, Af, rAU, patchPType, pBC,
                        phiStar, pCorr, phiNo, phi);
        or_continuity(nC, nF, nIF, owner, neigh, phi, divCorr);
      }
    }
    double mr = 0.0;
    for (int c = 0; c < nC; ++c) mr = fmax(mr, fabs(divCorr[c]));
    st.massRes = mr;
    for (int c = 0; c < nC; ++c) p[c] += par->pRelax * pCorr[c];
    or_lsq_gradient_gsv1(nC, nF, nIF, owner, ro, ci, cc, xf, tP, bp, pCorr, g);
    for (int c = 0; c < nC; ++c) {
      u[c] = u[c] - rAU[c] * g[3 * c];
      v[c] = v[c] - rAU[c] * g[3 * c + 1];
      w[c] = w[c] - rAU[c] * g[3 * c + 2];
    }
    st.duRel = relchg(nC, u, qOld[0]);
    st.dvRel = relchg(nC, v, qOld[1]);
    st.dwRel = relchg(nC, w, qOld[2]);
    st.dpRel = relchg(nC, p, pOld);
    if (stats) stats[step - 1] = st;
    done = step;
    if (!isfinite(mr) || mr > 1e50) {
      diverged = 1;
      break;
    }
    double mx = fmax(st.duRel, fmax(st.dvRel, st.dwRel));
    if (mr < par->tolMass && mx < par->tolVel) break;
  }
  if (phiOut) memcpy(phiOut, phi, sizeof(double) * (size_t)nF);
  if (rAUout) memcpy(rAUout, rAU, sizeof(double) * (size_t)nC);
  or_pcg_destroy(ps);
  free(tU);
  free(tP);
  free(bp);
  for (int k = 0; k < 3; ++k) {
    free(bv[k]);
    free(qOld[k]);
    free(gp[k]);
  }
  free(g);
  free(gpi);
  free(pOld);
  free(pCorr);
  free(rAU);
  free(qNew);
  free(vals);
  free(pv);
  free(rhs);
  free(phiStar);
  free(phi);
  free(phiNo);
  free(divStar);
  free(divCorr);
  free(divNo);
  return diverged ? -done : done;
}

/* threads the parallel loops use (OMP_NUM_THREADS or all cores) */
int or_num_threads(void) { return omp_get_max_threads(); }
