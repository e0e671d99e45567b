/*
 * simple_ops.c -- oracle restatement of the SIMPLE face/cell operators.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 *   or_rhiechow_flux     GSV1 main.cu:1474-1521
 *   or_continuity        GSV1 main.cu:1523-1536 (serial face order)
 *   or_pressure_nonorth  GSV1 main.cu:1538-1593 (serial face order)
 *   or_correct_flux      GSV1 main.cu:1615-1658
 *   or_patch_forces      GSV1 main.cu:2828-2933
 *
 * Pinned by tests/test_oracle_ops.py: with uniform rAU and a linear
 * pressure field the Rhie-Chow term vanishes (main.cu:694 "linear-pressure-
 * preserving"), so phi* equals rho U.S for a uniform velocity, and a
 * non-linear pressure switches it on; the continuity residual is zero for a
 * uniform velocity, sums to the net boundary outflow, and adds +phi to the
 * owner and -phi to the neighbour; for a linear p' and uniform rAU the
 * compact correction plus the non-orthogonal flux equals rAU grad(p').S on
 * internal and Dirichlet-p faces, Neumann faces keep phi*, and the
 * non-orthogonal flux vanishes on orthogonal meshes; the wall force of a
 * uniform pressure over a closed surface is zero, on a flat patch it is
 * p A n, the viscous force of a uniform tangential velocity is mu U A/dn,
 * and the pressure force is linear in p.  And by tests/test_oracle_pins.py:
 * the Rhie-Chow flux has the closed forms rho A U.n - rho A rAU dp / h (zero
 * cell gradients) and rho A U.n + rho A rAU a h^2 / 2 (cubic pressure with
 * exact cell gradients, so the sign of the term is fixed); the correction
 * and continuity operators run inside the plane-Poiseuille and
 * driver-restatement checks of the SIMPLE loop.
 */
#include "oracle.h"

#include <math.h>
#include <string.h>

void or_rhiechow_flux(int nFaces, int nInternalFaces, const int *owner,
                      const int *neigh, const int *bPatch, const double *cc,
                      const double *xf, const double *nf, const double *Af,
                      const double *u, const double *v, const double *w,
                      const double *p, const double *gradP,
                      const double *rAU, const int *bcUType,
                      const double *uFaceBC, const int *bcVType,
                      const double *vFaceBC, const int *bcWType,
                      const double *wFaceBC, double rho, double *phiStar) {
#pragma omp parallel for schedule(static)
  for (int f = 0; f < nFaces; ++f) {
    const int P = owner[f];
    if (f < nInternalFaces) {
      const int N = neigh[f];
      const double dx = cc[3 * N] - cc[3 * P];
      const double dy = cc[3 * N + 1] - cc[3 * P + 1];
      const double dz = cc[3 * N + 2] - cc[3 * P + 2];
      const double dpn = nf[3 * f] * dx + nf[3 * f + 1] * dy + nf[3 * f + 2] * dz;
      const double denom = fmax(dx * dx + dy * dy + dz * dz, 1.0e-30);
      double lam = ((xf[3 * f] - cc[3 * P]) * dx + (xf[3 * f + 1] - cc[3 * P + 1]) * dy +
                    (xf[3 * f + 2] - cc[3 * P + 2]) * dz) /
                   denom;
      lam = fmin(1.0, fmax(0.0, lam));
      const double ubar = (1.0 - lam) * u[P] + lam * u[N];
      const double vbar = (1.0 - lam) * v[P] + lam * v[N];
      const double wbar = (1.0 - lam) * w[P] + lam * w[N];
      const double gpx = (1.0 - lam) * gradP[3 * P] + lam * gradP[3 * N];
      const double gpy = (1.0 - lam) * gradP[3 * P + 1] + lam * gradP[3 * N + 1];
      const double gpz = (1.0 - lam) * gradP[3 * P + 2] + lam * gradP[3 * N + 2];
      const double rAUf = (1.0 - lam) * rAU[P] + lam * rAU[N];
      const double phiInterp =
          rho * Af[f] * (ubar * nf[3 * f] + vbar * nf[3 * f + 1] + wbar * nf[3 * f + 2]);
      const double rc = rho * Af[f] * rAUf / fmax(fabs(dpn), 1.0e-30) *
                        ((p[N] - p[P]) - (gpx * dx + gpy * dy + gpz * dz));
      phiStar[f] = phiInterp - rc;
    } else {
      const int patch = bPatch[f] - 1;
      const double uf = (bcUType[patch] == 1) ? uFaceBC[f] : u[P];
      const double vf = (bcVType[patch] == 1) ? vFaceBC[f] : v[P];
      const double wf = (bcWType[patch] == 1) ? wFaceBC[f] : w[P];
      phiStar[f] = rho * Af[f] * (uf * nf[3 * f] + vf * nf[3 * f + 1] + wf * nf[3 * f + 2]);
    }
  }
}

void or_continuity(int nCells, int nFaces, int nInternalFaces,
                   const int *owner, const int *neigh, const double *phi,
                   double *div) {
  memset(div, 0, sizeof(double) * (size_t)nCells);
  for (int f = 0; f < nFaces; ++f) {
    div[owner[f]] += phi[f];
    if (f < nInternalFaces) div[neigh[f]] += -phi[f];
  }
}

void or_pressure_nonorth(int nCells, int nFaces, int nInternalFaces,
                         const int *owner, const int *neigh,
                         const int *bPatch, const double *cc,
                         const double *xf, const double *nf,
                         const double *Sf, const double *Af,
                         const double *rAU, const int *bcPType,
                         const double *gradP, double *phiNonOrth,
                         double *divNonOrth) {
  memset(divNonOrth, 0, sizeof(double) * (size_t)nCells);
  for (int f = 0; f < nFaces; ++f) {
    double flux = 0.0;
    if (f < nInternalFaces) {
      int P = owner[f], N = neigh[f];
      double dx = cc[3 * N] - cc[3 * P];
      double dy = cc[3 * N + 1] - cc[3 * P + 1];
      double dz = cc[3 * N + 2] - cc[3 * P + 2];
      double dpn = nf[3 * f] * dx + nf[3 * f + 1] * dy + nf[3 * f + 2] * dz;
      if (fabs(dpn) > 1.0e-30) {
        double denom = fmax(dx * dx + dy * dy + dz * dz, 1.0e-30);
        double lam = ((xf[3 * f] - cc[3 * P]) * dx + (xf[3 * f + 1] - cc[3 * P + 1]) * dy +
                      (xf[3 * f + 2] - cc[3 * P + 2]) * dz) /
                     denom;
        lam = fmin(1.0, fmax(0.0, lam));
        double rAUf = (1.0 - lam) * rAU[P] + lam * rAU[N];
        double gx = (1.0 - lam) * gradP[3 * P] + lam * gradP[3 * N];
        double gy = (1.0 - lam) * gradP[3 * P + 1] + lam * gradP[3 * N + 1];
        double gz = (1.0 - lam) * gradP[3 * P + 2] + lam * gradP[3 * N + 2];
        double tx = Sf[3 * f] - (Af[f] / dpn) * dx;
        double ty = Sf[3 * f + 1] - (Af[f] / dpn) * dy;
        double tz = Sf[3 * f + 2] - (Af[f] / dpn) * dz;
        flux = rAUf * (gx * tx + gy * ty + gz * tz);
        divNonOrth[P] += flux;
        divNonOrth[N] += -flux;
      }
    } else {
      int patch = bPatch[f] - 1;
      if (patch >= 0 && bcPType[patch] == 1) {
        int P = owner[f];
        double dx = xf[3 * f] - cc[3 * P];
        double dy = xf[3 * f + 1] - cc[3 * P + 1];
        double dz = xf[3 * f + 2] - cc[3 * P + 2];
        double dpn = nf[3 * f] * dx + nf[3 * f + 1] * dy + nf[3 * f + 2] * dz;
        if (fabs(dpn) > 1.0e-30) {
          double tx = Sf[3 * f] - (Af[f] / dpn) * dx;
          double ty = Sf[3 * f + 1] - (Af[f] / dpn) * dy;
          double tz = Sf[3 * f + 2] - (Af[f] / dpn) * dz;
          flux = rAU[P] * (gradP[3 * P] * tx + gradP[3 * P + 1] * ty + gradP[3 * P + 2] * tz);
          divNonOrth[P] += flux;
        }
      }
    }
    phiNonOrth[f] = flux;
  }
}

void or_correct_flux(int nFaces, int nInternalFaces, const int *owner,
                     const int *neigh, const int *bPatch, const double *cc,
                     const double *xf, const double *nf, const double *Af,
                     const double *rAU, const int *bcPType,
                     const double *pFaceBC, const double *phiStar,
                     const double *pCorr, const double *phiNonOrth,
                     double *phi) {
#pragma omp parallel for schedule(static)
  for (int f = 0; f < nFaces; ++f) {
    if (f < nInternalFaces) {
      int P = owner[f], N = neigh[f];
      double dx = cc[3 * N] - cc[3 * P];
      double dy = cc[3 * N + 1] - cc[3 * P + 1];
      double dz = cc[3 * N + 2] - cc[3 * P + 2];
      double dpn = nf[3 * f] * dx + nf[3 * f + 1] * dy + nf[3 * f + 2] * dz;
      double denom = fmax(dx * dx + dy * dy + dz * dz, 1.0e-30);
      double lam = ((xf[3 * f] - cc[3 * P]) * dx + (xf[3 * f + 1] - cc[3 * P + 1]) * dy +
                    (xf[3 * f + 2] - cc[3 * P + 2]) * dz) /
                   denom;
      lam = fmin(1.0, fmax(0.0, lam));
      double rAUf = (1.0 - lam) * rAU[P] + lam * rAU[N];
      double coeff = Af[f] * rAUf / dpn;
      phi[f] = phiStar[f] - coeff * (pCorr[N] - pCorr[P]) - phiNonOrth[f];
    } else {
      int patch = bPatch[f] - 1;
      if (patch >= 0 && bcPType[patch] == 1) {
        int P = owner[f];
        double dx = xf[3 * f] - cc[3 * P];
        double dy = xf[3 * f + 1] - cc[3 * P + 1];
        double dz = xf[3 * f + 2] - cc[3 * P + 2];
        double dpn = nf[3 * f] * dx + nf[3 * f + 1] * dy + nf[3 * f + 2] * dz;
        if (fabs(dpn) <= 1.0e-30) {
          phi[f] = phiStar[f];
          continue;
        }
        double coeff = Af[f] * rAU[P] / dpn;
        phi[f] = phiStar[f] - coeff * (pFaceBC[f] - pCorr[P]) - phiNonOrth[f];
      } else {
        phi[f] = phiStar[f];
      }
    }
  }
}

static double norm3(const double *a) {
  return sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
}

void or_patch_forces(const int *owner, const double *cc, const double *xf,
                     const double *nf, const double *Af, int f0, int f1,
                     const double *u, const double *v, const double *w,
                     const double *p, double mu, int normalSign,
                     double *out) {
  double Fp[3] = {0, 0, 0}, Fv[3] = {0, 0, 0}, F[3] = {0, 0, 0};
  double area = 0.0, minD = 1.0e300, maxD = 0.0, maxUt = 0.0, maxSh = 0.0;
  int nFacesUsed = 0;
  for (int f = f0; f < f1; ++f) {
    const int P = owner[f];
    const double A = Af[f];
    if (A <= 1.0e-300) continue;
    double n[3] = {(double)normalSign * nf[3 * f], (double)normalSign * nf[3 * f + 1],
                   (double)normalSign * nf[3 * f + 2]};
    const double nmag = norm3(n);
    if (nmag <= 1.0e-300) continue;
    double s = 1.0 / nmag;
    n[0] = s * n[0];
    n[1] = s * n[1];
    n[2] = s * n[2];
    double dx[3] = {cc[3 * P] - xf[3 * f], cc[3 * P + 1] - xf[3 * f + 1],
                    cc[3 * P + 2] - xf[3 * f + 2]};
    double dn = dx[0] * n[0] + dx[1] * n[1] + dx[2] * n[2];
    if (dn <= 1.0e-14) dn = fabs(dn);
    if (dn <= 1.0e-14) dn = norm3(dx);
    if (dn <= 1.0e-14) continue;
    const double U[3] = {u[P], v[P], w[P]};
    const double Un = U[0] * n[0] + U[1] * n[1] + U[2] * n[2];
    const double Ut[3] = {U[0] - Un * n[0], U[1] - Un * n[1], U[2] - Un * n[2]};
    const double Utmag = norm3(Ut);
    const double shearMag = mu * Utmag / dn;
    for (int d = 0; d < 3; ++d) {
      const double fp = -p[P] * n[d] * A;
      const double fv = mu * Ut[d] / dn * A;
      Fp[d] += fp;
      Fv[d] += fv;
      F[d] += fp + fv;
    }
    area += A;
    nFacesUsed += 1;
    if (dn < minD) minD = dn;
    if (dn > maxD) maxD = dn;
    if (Utmag > maxUt) maxUt = Utmag;
    if (shearMag > maxSh) maxSh = shearMag;
  }
  if (minD == 1.0e300) minD = 0.0;
  for (int d = 0; d < 3; ++d) {
    out[d] = Fp[d];
    out[3 + d] = Fv[d];
    out[6 + d] = F[d];
  }
  out[9] = area;
  out[10] = (double)nFacesUsed;
  out[11] = minD;
  out[12] = maxD;
  out[13] = maxUt;
  out[14] = maxSh;
}
