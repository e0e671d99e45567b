/*
 * assembly.c -- oracle restatement of the face-loop assemblies.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * All loops run serially in face order (internal faces, then boundary faces),
 * which is the order of the reference host code and of one thread of the
 * reference's atomicAdd kernels; the products are written with the
 * reference's operation order.
 *
 *   or_elliptic_assemble   LP/src/scalar_elliptic.cu:60-136
 *   or_transport_assemble  LS/src/scalar_transport_library.cu:250-357
 *   or_momentum_assemble   GSV1 main.cu:1673-1684 (zero, base), 1804-1849
 *                          (internal faces), 1851-1896 (boundary faces)
 *   or_momentum_rhs_only   GSV1 main.cu:1697-1785
 *   or_momentum_relax      GSV1 main.cu:1686-1695
 *   or_momentum_relax_rhs_only GSV1 main.cu:1787-1794
 *   or_extract_rau         GSV1 main.cu:1796-1802
 *   or_pressure_assemble   GSV1 main.cu:1898-1950 (+ zero, 2532-2551)
 *
 * Pinned by tests/test_oracle.py and tests/test_oracle_ops.py: on an
 * orthogonal hex box the elliptic coefficient equals gamma*A/h exactly and
 * rows of interior cells sum to zero; linear fields are reproduced exactly
 * with Dirichlet + Neumann data in both sign conventions (LP -gamma g A, LS
 * +gamma g A) and with the reference cell of a pure-Neumann problem; the
 * pressure matrix is symmetric positive definite, linear in rAU, and equals
 * the elliptic matrix when rAU = gamma = 1 on an orthogonal mesh; the
 * momentum matrix with mu = 0 equals the LS central-convection matrix of the
 * same face fluxes (matrix and rhs, boundary rules included); upwind
 * convection has its closed form; the non-orthogonal correction vanishes on
 * orthogonal meshes; a Dirichlet MMS solve converges at second order in L2;
 * relaxation keeps the solution of the unrelaxed system and rAU = V/aP;
 * the rhs-only path plus its relaxation (aP rebuilt from the relaxed matrix,
 * main.cu:1790) equals the full assembly's relaxed rhs to round-off.  And by
 * tests/test_oracle_pins.py: the momentum assembly with rho = 0 equals the
 * LP elliptic assembly with gamma = mu on a jittered box (viscous
 * coefficients, Dirichlet diagonal and rhs, non-orthogonal rhs and source);
 * central convection on a uniform orthogonal box has the closed form
 * +-U_k A_k / 2 off the diagonal, a zero interior diagonal, F_b on an
 * outflow boundary diagonal and -F_b q_B on a Dirichlet rhs (transport and
 * momentum alike); the plane-Poiseuille and driver-restatement checks of the
 * SIMPLE loop exercise the momentum and pressure assemblies end to end.
 */
#include "oracle.h"

#include <math.h>
#include <string.h>

static inline double dot3(const double *a, const double *b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

static inline double lam_of(const double *cc, const double *xf, int P, int N,
                            int f) {
  /* face_interp_lambda (scalar_elliptic.cu:5-12) */
  double d[3] = {cc[3 * N] - cc[3 * P], cc[3 * N + 1] - cc[3 * P + 1],
                 cc[3 * N + 2] - cc[3 * P + 2]};
  double dx[3] = {xf[3 * f] - cc[3 * P], xf[3 * f + 1] - cc[3 * P + 1],
                  xf[3 * f + 2] - cc[3 * P + 2]};
  double dd = dot3(d, d);
  double lam = dot3(dx, d) / (dd > 1e-30 ? dd : 1e-30);
  lam = lam > 0.0 ? lam : 0.0;
  return lam < 1.0 ? lam : 1.0;
}

int or_elliptic_assemble(int nCells, int nFaces, int nInternalFaces,
                         const int *owner, const int *neigh,
                         const int *rowOffsets, const int *diagPos,
                         const int *facePP, const int *facePN,
                         const int *faceNP, const int *faceNN,
                         const double *cc, const double *vol,
                         const double *xf, const double *Sf, const double *Af,
                         const double *gammaFace, const double *cellSource,
                         const int *bcTypeBF, const double *bcValBF,
                         const double *grad, int includeNonOrth,
                         int useReferenceCell, int refCell, double refValue,
                         double *values, double *rhs) {
  memset(values, 0, sizeof(double) * (size_t)rowOffsets[nCells]);
  for (int c = 0; c < nCells; ++c) rhs[c] = cellSource[c] * vol[c];

  for (int f = 0; f < nInternalFaces; ++f) {
    int P = owner[f], N = neigh[f];
    const double *S = Sf + 3 * (size_t)f;
    double d[3] = {cc[3 * N] - cc[3 * P], cc[3 * N + 1] - cc[3 * P + 1],
                   cc[3 * N + 2] - cc[3 * P + 2]};
    double dDotS = dot3(d, S);
    double D = gammaFace[f] * dot3(S, S) / (dDotS > 1e-30 ? dDotS : 1e-30);
    values[facePP[f]] += D;
    values[facePN[f]] -= D;
    values[faceNP[f]] -= D;
    values[faceNN[f]] += D;
    if (includeNonOrth) {
      double g = gammaFace[f];
      double s = D / (g > 1e-30 ? g : 1e-30);
      double T[3] = {S[0] - s * d[0], S[1] - s * d[1], S[2] - s * d[2]};
      double lam = lam_of(cc, xf, P, N, f);
      double gF[3];
      for (int q = 0; q < 3; ++q)
        gF[q] = (1.0 - lam) * grad[3 * P + q] + lam * grad[3 * N + q];
      double corr = g * dot3(T, gF);
      rhs[P] += corr;
      rhs[N] -= corr;
    }
  }
  for (int f = nInternalFaces; f < nFaces; ++f) {
    int P = owner[f];
    int bf = f - nInternalFaces;
    const double *S = Sf + 3 * (size_t)f;
    double d[3] = {xf[3 * f] - cc[3 * P], xf[3 * f + 1] - cc[3 * P + 1],
                   xf[3 * f + 2] - cc[3 * P + 2]};
    double dDotS = dot3(d, S);
    double D = gammaFace[f] * dot3(S, S) / (dDotS > 1e-30 ? dDotS : 1e-30);
    if (bcTypeBF[bf] == OR_BC_DIRICHLET) {
      values[diagPos[P]] += D;
      rhs[P] += D * bcValBF[bf];
      if (includeNonOrth) {
        double g = gammaFace[f];
        double s = D / (g > 1e-30 ? g : 1e-30);
        double T[3] = {S[0] - s * d[0], S[1] - s * d[1], S[2] - s * d[2]};
        double corr = g * dot3(T, grad + 3 * (size_t)P);
        rhs[P] += corr;
      }
    } else {
      rhs[P] += -gammaFace[f] * bcValBF[bf] * Af[f];
    }
  }
  if (useReferenceCell) {
    if (refCell < 0 || refCell >= nCells) return -1;
    for (int k = rowOffsets[refCell]; k < rowOffsets[refCell + 1]; ++k)
      values[k] = 0.0;
    values[diagPos[refCell]] = 1.0;
    rhs[refCell] = refValue;
  }
  return 0;
}

int or_transport_assemble(int nCells, int nFaces, int nInternalFaces,
                          const int *owner, const int *neigh,
                          const int *diagPos, const int *facePP,
                          const int *facePN, const int *faceNP,
                          const int *faceNN, const double *cc,
                          const double *vol, const double *xf,
                          const double *Sf, const double *Af,
                          const double *faceFlux, const double *gammaFace,
                          const double *Su, const double *Sp,
                          const int *bcTypeBF, const double *bcValBF,
                          const double *grad, int includeNonOrth, int upwind,
                          double *values, double *rhs) {
  int64_t nnz = 0;
  for (int c = 0; c < nCells; ++c)
    if (diagPos[c] + 1 > nnz) nnz = diagPos[c] + 1;
  for (int f = 0; f < nInternalFaces; ++f) {
    if (facePN[f] + 1 > nnz) nnz = facePN[f] + 1;
    if (faceNP[f] + 1 > nnz) nnz = faceNP[f] + 1;
  }
  memset(values, 0, sizeof(double) * (size_t)nnz);
  for (int c = 0; c < nCells; ++c) {
    rhs[c] = Su[c] * vol[c];
    values[diagPos[c]] += -Sp[c] * vol[c];
  }
  for (int f = 0; f < nInternalFaces; ++f) {
    int P = owner[f], N = neigh[f];
    const double *S = Sf + 3 * (size_t)f;
    double d[3] = {cc[3 * N] - cc[3 * P], cc[3 * N + 1] - cc[3 * P + 1],
                   cc[3 * N + 2] - cc[3 * P + 2]};
    double dDotS = dot3(d, S);
    double gamma = gammaFace[f];
    double D = gamma * dot3(S, S) / (dDotS > 1e-30 ? dDotS : 1e-30);
    values[facePP[f]] += D;
    values[facePN[f]] -= D;
    values[faceNP[f]] -= D;
    values[faceNN[f]] += D;
    if (includeNonOrth) {
      double s = D / (gamma > 1e-30 ? gamma : 1e-30);
      double T[3] = {S[0] - s * d[0], S[1] - s * d[1], S[2] - s * d[2]};
      double lam = lam_of(cc, xf, P, N, f);
      double gF[3];
      for (int q = 0; q < 3; ++q)
        gF[q] = (1.0 - lam) * grad[3 * P + q] + lam * grad[3 * N + q];
      double corr = gamma * dot3(T, gF);
      rhs[P] += corr;
      rhs[N] -= corr;
    }
    double F = faceFlux[f];
    if (!upwind) {
      double lam = lam_of(cc, xf, P, N, f);
      double aP = F * (1.0 - lam);
      double aN = F * lam;
      values[facePP[f]] += aP;
      values[facePN[f]] += aN;
      values[faceNP[f]] -= aP;
      values[faceNN[f]] -= aN;
    } else {
      if (F >= 0.0) {
        values[facePP[f]] += F;
        values[faceNP[f]] -= F;
      } else {
        values[facePN[f]] += F;
        values[faceNN[f]] -= F;
      }
    }
  }
  for (int f = nInternalFaces; f < nFaces; ++f) {
    int P = owner[f];
    int bf = f - nInternalFaces;
    const double *S = Sf + 3 * (size_t)f;
    double d[3] = {xf[3 * f] - cc[3 * P], xf[3 * f + 1] - cc[3 * P + 1],
                   xf[3 * f + 2] - cc[3 * P + 2]};
    double dDotS = dot3(d, S);
    double gamma = gammaFace[f];
    double D = gamma * dot3(S, S) / (dDotS > 1e-30 ? dDotS : 1e-30);
    double F = faceFlux[f];
    if (bcTypeBF[bf] == OR_BC_DIRICHLET) {
      double phiB = bcValBF[bf];
      values[diagPos[P]] += D;
      rhs[P] += D * phiB;
      if (includeNonOrth) {
        double s = D / (gamma > 1e-30 ? gamma : 1e-30);
        double T[3] = {S[0] - s * d[0], S[1] - s * d[1], S[2] - s * d[2]};
        double corr = gamma * dot3(T, grad + 3 * (size_t)P);
        rhs[P] += corr;
      }
      rhs[P] += -F * phiB;
    } else {
      double gradn = (bcTypeBF[bf] == OR_BC_ZEROGRAD) ? 0.0 : bcValBF[bf];
      rhs[P] += gamma * gradn * Af[f];
      values[diagPos[P]] += F;
    }
  }
  return 0;
}

/* kernel_momentum_internal_faces geometry terms (main.cu:1818-1842) */
typedef struct {
  double alpha, tcx, tcy, tcz, lam, dx, dy, dz, nd;
} mom_face;

static int mom_internal(const double *cc, const double *xf, const double *nf,
                        const double *Sf, const double *Af, int P, int N,
                        int f, double mu, mom_face *m) {
  double dx = cc[3 * N] - cc[3 * P];
  double dy = cc[3 * N + 1] - cc[3 * P + 1];
  double dz = cc[3 * N + 2] - cc[3 * P + 2];
  double nd = nf[3 * f] * dx + nf[3 * f + 1] * dy + nf[3 * f + 2] * dz;
  if (nd <= 1e-14) return 0;
  double af = Af[f];
  m->alpha = mu * af / nd;
  m->tcx = mu * (Sf[3 * f] - (af / nd) * dx);
  m->tcy = mu * (Sf[3 * f + 1] - (af / nd) * dy);
  m->tcz = mu * (Sf[3 * f + 2] - (af / nd) * dz);
  double denom = dx * dx + dy * dy + dz * dz;
  double lam = ((xf[3 * f] - cc[3 * P]) * dx + (xf[3 * f + 1] - cc[3 * P + 1]) * dy +
                (xf[3 * f + 2] - cc[3 * P + 2]) * dz) /
               (denom > 1e-30 ? denom : 1e-30);
  lam = fmin(1.0, fmax(0.0, lam));
  m->lam = lam;
  m->dx = dx;
  m->dy = dy;
  m->dz = dz;
  m->nd = nd;
  return 1;
}

static int mom_boundary(const double *cc, const double *xf, const double *nf,
                        const double *Sf, const double *Af, int P, int f,
                        double mu, mom_face *m) {
  double dx = xf[3 * f] - cc[3 * P];
  double dy = xf[3 * f + 1] - cc[3 * P + 1];
  double dz = xf[3 * f + 2] - cc[3 * P + 2];
  double nd = nf[3 * f] * dx + nf[3 * f + 1] * dy + nf[3 * f + 2] * dz;
  if (nd <= 1e-14) return 0;
  double af = Af[f];
  m->alpha = mu * af / nd;
  m->tcx = mu * (Sf[3 * f] - (af / nd) * dx);
  m->tcy = mu * (Sf[3 * f + 1] - (af / nd) * dy);
  m->tcz = mu * (Sf[3 * f + 2] - (af / nd) * dz);
  return 1;
}

void or_momentum_assemble(int nCells, int nFaces, int nInternalFaces,
                          const int *owner, const int *neigh,
                          const int *bPatch, const int *diagPos,
                          const int *facePP, const int *facePN,
                          const int *faceNP, const int *faceNN,
                          const double *cc, const double *vol,
                          const double *xf, const double *nf,
                          const double *Sf, const double *Af,
                          const double *gradQ, const double *uConv,
                          const double *vConv, const double *wConv,
                          const double *gradPcomp, const int *bcQType,
                          const double *bcQVal, const int *bcUType,
                          const double *bcUVal, const int *bcVType,
                          const double *bcVVal, const int *bcWType,
                          const double *bcWVal, double rho, double mu,
                          double corrPsi, int64_t nnz, double *vals,
                          double *rhs) {
  memset(vals, 0, sizeof(double) * (size_t)nnz);
  for (int c = 0; c < nCells; ++c) {
    vals[diagPos[c]] = 0.0;
    rhs[c] = -gradPcomp[c] * vol[c];
  }
  for (int f = 0; f < nInternalFaces; ++f) {
    int P = owner[f], N = neigh[f];
    mom_face m;
    if (!mom_internal(cc, xf, nf, Sf, Af, P, N, f, mu, &m)) continue;
    double lam = m.lam;
    double gx = (1.0 - lam) * gradQ[3 * P] + lam * gradQ[3 * N];
    double gy = (1.0 - lam) * gradQ[3 * P + 1] + lam * gradQ[3 * N + 1];
    double gz = (1.0 - lam) * gradQ[3 * P + 2] + lam * gradQ[3 * N + 2];
    double corr = corrPsi * (gx * m.tcx + gy * m.tcy + gz * m.tcz);
    double ucf = (1.0 - lam) * uConv[P] + lam * uConv[N];
    double vcf = (1.0 - lam) * vConv[P] + lam * vConv[N];
    double wcf = (1.0 - lam) * wConv[P] + lam * wConv[N];
    double F = rho * Af[f] *
               (ucf * nf[3 * f] + vcf * nf[3 * f + 1] + wcf * nf[3 * f + 2]);
    vals[facePP[f]] += (m.alpha + F * (1.0 - lam));
    vals[facePN[f]] += (-m.alpha + F * lam);
    vals[faceNP[f]] += (-m.alpha - F * (1.0 - lam));
    vals[faceNN[f]] += (m.alpha - F * lam);
    rhs[P] += corr;
    rhs[N] += (-corr);
  }
  for (int f = nInternalFaces; f < nFaces; ++f) {
    int P = owner[f];
    int patch = bPatch[f] - 1;
    mom_face m;
    if (!mom_boundary(cc, xf, nf, Sf, Af, P, f, mu, &m)) continue;
    int diag = diagPos[P];
    if (bcQType[patch] == 1) {
      double g = bcQVal[f];
      double corr = corrPsi * (gradQ[3 * P] * m.tcx + gradQ[3 * P + 1] * m.tcy +
                               gradQ[3 * P + 2] * m.tcz);
      vals[diag] += m.alpha;
      rhs[P] += (m.alpha * g + corr);
    }
    double ucb = (bcUType[patch] == 1) ? bcUVal[f] : uConv[P];
    double vcb = (bcVType[patch] == 1) ? bcVVal[f] : vConv[P];
    double wcb = (bcWType[patch] == 1) ? bcWVal[f] : wConv[P];
    double F = rho * Af[f] *
               (ucb * nf[3 * f] + vcb * nf[3 * f + 1] + wcb * nf[3 * f + 2]);
    if (bcQType[patch] == 1)
      rhs[P] += (-F * bcQVal[f]);
    else
      vals[diag] += F;
  }
}

void or_momentum_rhs_only(int nCells, int nFaces, int nInternalFaces,
                          const int *owner, const int *neigh,
                          const int *bPatch, const double *cc,
                          const double *vol, const double *xf,
                          const double *nf, const double *Sf,
                          const double *Af, const double *gradQ,
                          const double *uConv, const double *vConv,
                          const double *wConv, const double *gradPcomp,
                          const int *bcQType, const double *bcQVal,
                          const int *bcUType, const double *bcUVal,
                          const int *bcVType, const double *bcVVal,
                          const int *bcWType, const double *bcWVal,
                          double rho, double mu, double corrPsi,
                          double *rhs) {
  for (int c = 0; c < nCells; ++c) rhs[c] = -gradPcomp[c] * vol[c];
  for (int f = 0; f < nInternalFaces; ++f) {
    int P = owner[f], N = neigh[f];
    mom_face m;
    if (!mom_internal(cc, xf, nf, Sf, Af, P, N, f, mu, &m)) continue;
    double lam = m.lam;
    double gx = (1.0 - lam) * gradQ[3 * P] + lam * gradQ[3 * N];
    double gy = (1.0 - lam) * gradQ[3 * P + 1] + lam * gradQ[3 * N + 1];
    double gz = (1.0 - lam) * gradQ[3 * P + 2] + lam * gradQ[3 * N + 2];
    double corr = corrPsi * (gx * m.tcx + gy * m.tcy + gz * m.tcz);
    rhs[P] += corr;
    rhs[N] += (-corr);
  }
  for (int f = nInternalFaces; f < nFaces; ++f) {
    int P = owner[f];
    int patch = bPatch[f] - 1;
    mom_face m;
    if (!mom_boundary(cc, xf, nf, Sf, Af, P, f, mu, &m)) continue;
    if (bcQType[patch] == 1) {
      double g = bcQVal[f];
      double corr = corrPsi * (gradQ[3 * P] * m.tcx + gradQ[3 * P + 1] * m.tcy +
                               gradQ[3 * P + 2] * m.tcz);
      rhs[P] += (m.alpha * g + corr);
    }
    double ucb = (bcUType[patch] == 1) ? bcUVal[f] : uConv[P];
    double vcb = (bcVType[patch] == 1) ? bcVVal[f] : vConv[P];
    double wcb = (bcWType[patch] == 1) ? bcWVal[f] : wConv[P];
    double F = rho * Af[f] *
               (ucb * nf[3 * f] + vcb * nf[3 * f + 1] + wcb * nf[3 * f + 2]);
    if (bcQType[patch] == 1) rhs[P] += (-F * bcQVal[f]);
  }
}

void or_momentum_relax(int nCells, const int *diagPos, double *vals,
                       double *rhs, const double *qOld, double uRelax) {
  if (uRelax >= 0.999999) return; /* relax_momentum_system_on_gpu :2522 */
  for (int c = 0; c < nCells; ++c) {
    int diag = diagPos[c];
    double aP = vals[diag];
    double invRelax = 1.0 / uRelax;
    vals[diag] = aP * invRelax;
    rhs[c] += (invRelax - 1.0) * aP * qOld[c];
  }
}

void or_momentum_relax_rhs_only(int nCells, const int *diagPos,
                                const double *vals, double *rhs,
                                const double *qOld, double uRelax) {
  if (!(uRelax < 0.999999)) return; /* main.cu:2502 */
  for (int c = 0; c < nCells; ++c) {
    double aP_old = vals[diagPos[c]] * uRelax;
    double invRelax = 1.0 / uRelax;
    rhs[c] += (invRelax - 1.0) * aP_old * qOld[c];
  }
}

void or_extract_rau(int nCells, const int *diagPos, const double *vals,
                    const double *vol, double *rAU) {
  for (int c = 0; c < nCells; ++c) {
    double a = vals[diagPos[c]];
    rAU[c] = (fabs(a) > 1e-30) ? vol[c] / a : 0.0;
  }
}

void or_pressure_assemble(int nCells, int nFaces, int nInternalFaces,
                          const int *owner, const int *neigh,
                          const int *bPatch, const int *diagPos,
                          const int *facePP, const int *facePN,
                          const int *faceNP, const int *faceNN,
                          const double *cc, const double *xf,
                          const double *nf, const double *Af,
                          const double *rAU, const int *bcPType,
                          int64_t nnz, double *vals) {
  (void)nCells;
  memset(vals, 0, sizeof(double) * (size_t)nnz);
  for (int f = 0; f < nInternalFaces; ++f) {
    int P = owner[f], N = neigh[f];
    double dx = cc[3 * N] - cc[3 * P];
    double dy = cc[3 * N + 1] - cc[3 * P + 1];
    double dz = cc[3 * N + 2] - cc[3 * P + 2];
    double dpn = nf[3 * f] * dx + nf[3 * f + 1] * dy + nf[3 * f + 2] * dz;
    if (dpn <= 1e-14) continue;
    double denom = dx * dx + dy * dy + dz * dz;
    double lam = ((xf[3 * f] - cc[3 * P]) * dx + (xf[3 * f + 1] - cc[3 * P + 1]) * dy +
                  (xf[3 * f + 2] - cc[3 * P + 2]) * dz) /
                 (denom > 1e-30 ? denom : 1e-30);
    lam = fmin(1.0, fmax(0.0, lam));
    double rAUf = (1.0 - lam) * rAU[P] + lam * rAU[N];
    double coeff = Af[f] * rAUf / dpn;
    vals[facePP[f]] += coeff;
    vals[facePN[f]] += (-coeff);
    vals[faceNP[f]] += (-coeff);
    vals[faceNN[f]] += coeff;
  }
  for (int f = nInternalFaces; f < nFaces; ++f) {
    int patch = bPatch[f] - 1;
    if (bcPType[patch] != 1) continue;
    int P = owner[f];
    double dx = xf[3 * f] - cc[3 * P];
    double dy = xf[3 * f + 1] - cc[3 * P + 1];
    double dz = xf[3 * f + 2] - cc[3 * P + 2];
    double dpn = nf[3 * f] * dx + nf[3 * f + 1] * dy + nf[3 * f + 2] * dz;
    if (dpn <= 1e-14) continue;
    double coeff = Af[f] * rAU[P] / dpn;
    vals[diagPos[P]] += coeff;
  }
}
