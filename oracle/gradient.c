/*
 * gradient.c -- oracle restatement of the least-squares cell gradients.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * or_lsq_gradient_gsv1 follows GSV1 build_lsq_gradient_operator
 *   (apps/generic_simple_v1/src/main.cu:1159-1251) and kernel_apply_lsq_gradient
 *   (main.cu:1403-1440): terms = sorted unique interior neighbours, then the
 *   cell's boundary faces in face order; r = x_nb - x_P, w = 1/max(|r|^2,1e-30),
 *   explicit adjugate inverse when |det| > 1e-20 (else zero operator);
 *   c_t = M^-1 (w r_t); g = sum_t c_t (val_t - phi_P), with val = phi_P on
 *   non-Dirichlet boundary faces.
 * or_lsq_gradient_lp follows LP/src/gradient.cu:21-83 (Neumann faces use the
 *   normal constraint r = dn*n, dphi = g*dn; the sums are accumulated
 *   directly: g = M^-1 sum w r dphi).
 *
 * Pinned by tests/test_oracle.py (test_gradient_*): exact reproduction of linear fields
 * when every boundary face is an exact Dirichlet face (both variants), zero
 * gradient of constants, a brute-force numpy least-squares solve per cell on a
 * small jittered mesh (independent of the adjugate formula), and the LP
 * Neumann constraint reproducing a linear field whose normal derivative is
 * supplied exactly.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>

/* boundary faces of each cell, in face order (cellBFace, mesh.cu:262-273) */
static void cell_bfaces(int nCells, int nFaces, int nInternalFaces,
                        const int *owner, int **off_out, int **lst_out) {
  int *off = (int *)calloc((size_t)nCells + 1, sizeof(int));
  for (int f = nInternalFaces; f < nFaces; ++f) off[owner[f] + 1]++;
  for (int c = 0; c < nCells; ++c) off[c + 1] += off[c];
  int *lst = (int *)malloc(sizeof(int) * ((size_t)off[nCells] + 1));
  int *fill = (int *)malloc(sizeof(int) * (size_t)nCells);
  for (int c = 0; c < nCells; ++c) fill[c] = off[c];
  for (int f = nInternalFaces; f < nFaces; ++f) lst[fill[owner[f]]++] = f;
  free(fill);
  *off_out = off;
  *lst_out = lst;
}

static void add_term(double M[3][3], const double r[3], double w) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) M[i][j] += w * r[i] * r[j];
}

/* returns 1 and fills inv if |det| > 1e-20 (main.cu:1221-1232) */
static int adjugate_inverse(double M[3][3], double inv[3][3]) {
  double a = M[0][0], b = M[0][1], c = M[0][2];
  double d = M[1][0], e = M[1][1], f2 = M[1][2];
  double g = M[2][0], h = M[2][1], k = M[2][2];
  double det = a * (e * k - f2 * h) - b * (d * k - f2 * g) + c * (d * h - e * g);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) inv[i][j] = 0.0;
  if (!(fabs(det) > 1e-20)) return 0;
  inv[0][0] = (e * k - f2 * h) / det;
  inv[0][1] = (c * h - b * k) / det;
  inv[0][2] = (b * f2 - c * e) / det;
  inv[1][0] = (f2 * g - d * k) / det;
  inv[1][1] = (a * k - c * g) / det;
  inv[1][2] = (c * d - a * f2) / det;
  inv[2][0] = (d * h - e * g) / det;
  inv[2][1] = (b * g - a * h) / det;
  inv[2][2] = (a * e - b * d) / det;
  return 1;
}

void or_lsq_gradient_gsv1(int nCells, int nFaces, int nInternalFaces,
                          const int *owner, const int *rowOffsets,
                          const int *cols, const double *cc, const double *xf,
                          const int *bcTypeBF, const double *bcValBF,
                          const double *phi, double *grad) {
  int *bo, *bl;
  cell_bfaces(nCells, nFaces, nInternalFaces, owner, &bo, &bl);
#pragma omp parallel for schedule(static)
  for (int P = 0; P < nCells; ++P) {
    const double *xP = cc + 3 * (size_t)P;
    double M[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    /* pass 1: normal matrix */
    for (int k = rowOffsets[P]; k < rowOffsets[P + 1]; ++k) {
      int N = cols[k];
      if (N == P) continue;
      const double *xN = cc + 3 * (size_t)N;
      double r[3] = {xN[0] - xP[0], xN[1] - xP[1], xN[2] - xP[2]};
      double rr = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
      double w = 1.0 / (rr > 1e-30 ? rr : 1e-30);
      add_term(M, r, w);
    }
    for (int j = bo[P]; j < bo[P + 1]; ++j) {
      int f = bl[j];
      const double *x = xf + 3 * (size_t)f;
      double r[3] = {x[0] - xP[0], x[1] - xP[1], x[2] - xP[2]};
      double rr = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
      double w = 1.0 / (rr > 1e-30 ? rr : 1e-30);
      add_term(M, r, w);
    }
    double inv[3][3];
    adjugate_inverse(M, inv);
    /* pass 2: operator coefficients c_t and application, term order */
    const double phiP = phi[P];
    double sx = 0.0, sy = 0.0, sz = 0.0;
    for (int k = rowOffsets[P]; k < rowOffsets[P + 1]; ++k) {
      int N = cols[k];
      if (N == P) continue;
      const double *xN = cc + 3 * (size_t)N;
      double r[3] = {xN[0] - xP[0], xN[1] - xP[1], xN[2] - xP[2]};
      double rr = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
      double w = 1.0 / (rr > 1e-30 ? rr : 1e-30);
      double bx = w * r[0], by = w * r[1], bz = w * r[2];
      double cx = inv[0][0] * bx + inv[0][1] * by + inv[0][2] * bz;
      double cy = inv[1][0] * bx + inv[1][1] * by + inv[1][2] * bz;
      double cz = inv[2][0] * bx + inv[2][1] * by + inv[2][2] * bz;
      double dphi = phi[N] - phiP;
      sx += cx * dphi;
      sy += cy * dphi;
      sz += cz * dphi;
    }
    for (int j = bo[P]; j < bo[P + 1]; ++j) {
      int f = bl[j];
      const double *x = xf + 3 * (size_t)f;
      double r[3] = {x[0] - xP[0], x[1] - xP[1], x[2] - xP[2]};
      double rr = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
      double w = 1.0 / (rr > 1e-30 ? rr : 1e-30);
      double bx = w * r[0], by = w * r[1], bz = w * r[2];
      double cx = inv[0][0] * bx + inv[0][1] * by + inv[0][2] * bz;
      double cy = inv[1][0] * bx + inv[1][1] * by + inv[1][2] * bz;
      double cz = inv[2][0] * bx + inv[2][1] * by + inv[2][2] * bz;
      int bf = f - nInternalFaces;
      double val = (bcTypeBF[bf] == OR_BC_DIRICHLET) ? bcValBF[bf] : phiP;
      double dphi = val - phiP;
      sx += cx * dphi;
      sy += cy * dphi;
      sz += cz * dphi;
    }
    grad[3 * (size_t)P + 0] = sx;
    grad[3 * (size_t)P + 1] = sy;
    grad[3 * (size_t)P + 2] = sz;
  }
  free(bo);
  free(bl);
}

void or_lsq_gradient_lp(int nCells, int nFaces, int nInternalFaces,
                        const int *owner, const int *rowOffsets,
                        const int *cols, const double *cc, const double *xf,
                        const double *nf, const int *bcTypeBF,
                        const double *bcValBF, const double *phi,
                        double *grad) {
  int *bo, *bl;
  cell_bfaces(nCells, nFaces, nInternalFaces, owner, &bo, &bl);
#pragma omp parallel for schedule(static)
  for (int P = 0; P < nCells; ++P) {
    const double *xP = cc + 3 * (size_t)P;
    const double phiP = phi[P];
    double M[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    double rhs[3] = {0, 0, 0};
    for (int k = rowOffsets[P]; k < rowOffsets[P + 1]; ++k) {
      int N = cols[k];
      if (N == P) continue;
      const double *xN = cc + 3 * (size_t)N;
      double r[3] = {xN[0] - xP[0], xN[1] - xP[1], xN[2] - xP[2]};
      double dphi = phi[N] - phiP;
      double rr = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
      double w = 1.0 / (rr > 1e-30 ? rr : 1e-30);
      for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) M[i][j] += w * r[i] * r[j];
        rhs[i] += w * r[i] * dphi;
      }
    }
    for (int j = bo[P]; j < bo[P + 1]; ++j) {
      int f = bl[j];
      const double *x = xf + 3 * (size_t)f;
      double rcf[3] = {x[0] - xP[0], x[1] - xP[1], x[2] - xP[2]};
      double r[3], dphi;
      int bf = f - nInternalFaces;
      if (bcTypeBF[bf] == OR_BC_DIRICHLET) {
        r[0] = rcf[0];
        r[1] = rcf[1];
        r[2] = rcf[2];
        dphi = bcValBF[bf] - phiP;
      } else {
        const double *n = nf + 3 * (size_t)f;
        double dn = rcf[0] * n[0] + rcf[1] * n[1] + rcf[2] * n[2];
        dn = dn > 1e-30 ? dn : 1e-30;
        r[0] = dn * n[0];
        r[1] = dn * n[1];
        r[2] = dn * n[2];
        /* LS ZeroGradient carries value 0 (scalar_transport_library.cu:88) */
        dphi = bcValBF[bf] * dn;
      }
      double rr = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
      double w = 1.0 / (rr > 1e-30 ? rr : 1e-30);
      for (int i = 0; i < 3; ++i) {
        for (int jj = 0; jj < 3; ++jj) M[i][jj] += w * r[i] * r[jj];
        rhs[i] += w * r[i] * dphi;
      }
    }
    double inv[3][3];
    double g[3] = {0, 0, 0};
    if (adjugate_inverse(M, inv)) {
      g[0] = inv[0][0] * rhs[0] + inv[0][1] * rhs[1] + inv[0][2] * rhs[2];
      g[1] = inv[1][0] * rhs[0] + inv[1][1] * rhs[1] + inv[1][2] * rhs[2];
      g[2] = inv[2][0] * rhs[0] + inv[2][1] * rhs[1] + inv[2][2] * rhs[2];
    }
    grad[3 * (size_t)P + 0] = g[0];
    grad[3 * (size_t)P + 1] = g[1];
    grad[3 * (size_t)P + 2] = g[2];
  }
  free(bo);
  free(bl);
}
