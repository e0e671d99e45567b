// kernels.hpp -- device views and kernel launchers of the anabasis library.
#pragma once

#include "an_internal.hpp"

namespace an {

// POD view of a DevMesh passed by value to kernels.
struct MeshView {
  int nC, nF, nIF, nBF;
  const int *rowOff, *cols, *diagPos;
  const int *cfOff, *cfEnt, *cfSlot, *cbOff, *cbFace;
  const double4* cgeo;
  const int2* fPN;
  const double *fLam, *fDpn, *fA, *fNx, *fNy, *fNz, *fDx, *fDy, *fDz;
  const int *bP, *bPatch;
  const double *bDpn, *bA, *bNx, *bNy, *bNz, *bDx, *bDy, *bDz;
  const double *fTx, *fTy, *fTz, *bTx, *bTy, *bTz;  // T = A n - (A / n.d) d
  const double* gsvInv;  // GSV1 LSQ normal-matrix inverse, 9 x nC SoA (may be null)
  // GSV1 gradient as a sliced per-term operator (may be null): cells in
  // slices of 64, term j of cell c at lsqOff[c >> 6] + 64 j + (c & 63)
  const int *lsqOff, *lsqLen, *lsqSrc;
  const double *lsqCx, *lsqCy, *lsqCz;
};
MeshView view_of(const DevMesh& m);
// the GSV1 gradient's per-cell normal-matrix inverse (geometry only)
void build_gsv1_inverse(const MeshView& m, DevBuf<double>& inv9, hipStream_t s);
// the sliced GSV1 operator from the cached inverse (needs m.gsvInv)
void build_lsq_operator(const MeshView& m, LsqOperator& op, hipStream_t s);

// ------------------------------------------------ shared device helpers
// LSQ normal matrix and its inverse (gradient.hip, adjoint.hip)
__device__ __forceinline__ void acc_M(double M[9], double r0, double r1, double r2, double w) {
  // M[i][j] += w*r[i]*r[j] as the reference writes it: (w*r_i)*r_j is not
  // bitwise symmetric, so all nine entries are accumulated.
  const double a0 = w * r0, a1 = w * r1, a2 = w * r2;
  M[0] += a0 * r0;
  M[1] += a0 * r1;
  M[2] += a0 * r2;
  M[3] += a1 * r0;
  M[4] += a1 * r1;
  M[5] += a1 * r2;
  M[6] += a2 * r0;
  M[7] += a2 * r1;
  M[8] += a2 * r2;
}

// explicit adjugate inverse (main.cu:1221-1232); returns false (zero
// operator) when |det| <= 1e-20.
__device__ __forceinline__ bool inv3(const double M[9], double inv[9]) {
  const double a = M[0], b = M[1], c = M[2];
  const double d = M[3], e = M[4], f2 = M[5];
  const double g = M[6], h = M[7], k = M[8];
  const double det = a * (e * k - f2 * h) - b * (d * k - f2 * g) + c * (d * h - e * g);
  if (!(fabs(det) > 1e-20)) {
    for (int i = 0; i < 9; ++i) inv[i] = 0.0;
    return false;
  }
  inv[0] = (e * k - f2 * h) / det;
  inv[1] = (c * h - b * k) / det;
  inv[2] = (b * f2 - c * e) / det;
  inv[3] = (f2 * g - d * k) / det;
  inv[4] = (a * k - c * g) / det;
  inv[5] = (c * d - a * f2) / det;
  inv[6] = (d * h - e * g) / det;
  inv[7] = (b * g - a * h) / det;
  inv[8] = (a * e - b * d) / det;
  return true;
}

// face diffusion coefficient and non-orthogonal term (assembly.hip, adjoint.hip)
// D = gamma |S|^2 / max(d.S, 1e-30) with S = A n (Sf = mul3(areaMag, nloc),
// LP/src/mesh.cu:259) and d the owner->neighbour (or owner->face) vector.
struct DiffFace {
  double S0, S1, S2, d0, d1, d2, D;
};
__device__ __forceinline__ DiffFace diff_face(double A, double n0, double n1, double n2, double d0,
                                              double d1, double d2, double gamma) {
  DiffFace r;
  r.S0 = A * n0;
  r.S1 = A * n1;
  r.S2 = A * n2;
  r.d0 = d0;
  r.d1 = d1;
  r.d2 = d2;
  const double dDotS = d0 * r.S0 + d1 * r.S1 + d2 * r.S2;
  const double SS = r.S0 * r.S0 + r.S1 * r.S1 + r.S2 * r.S2;
  r.D = gamma * SS / (dDotS > 1e-30 ? dDotS : 1e-30);
  return r;
}

// non-orthogonal correction gamma * T.g with T = S - (D/max(gamma,1e-30)) d
__device__ __forceinline__ double nonorth_corr(const DiffFace& q, double gamma, double g0, double g1,
                                               double g2) {
  const double s = q.D / (gamma > 1e-30 ? gamma : 1e-30);
  const double T0 = q.S0 - s * q.d0, T1 = q.S1 - s * q.d1, T2 = q.S2 - s * q.d2;
  return gamma * (T0 * g0 + T1 * g1 + T2 * g2);
}

// Per-boundary-face BC (type AN_BC_*, value), device arrays of nBF.
struct BCView {
  const int* type;
  const double* val;
};

// ---------------------------------------------------------- gradients
// mode 0: GSV1 operator semantics, 1: LP compute_lsq_gradient.  Output SoA.
void launch_lsq_gradient(const MeshView& m, int mode, BCView bc, const double* phi, double* gx,
                         double* gy, double* gz, hipStream_t s);
// GSV1 gradients of three fields sharing one BC type array (u, v, w): the
// geometry, the cached inverse and the adjacency are read once; each field's
// sums are those of launch_lsq_gradient (mode 0), so the results are the
// same bits.  Needs the cached GSV1 inverse.
struct Grad3Args {
  const int* bcType;
  const double* bcVal[3];
  const double* phi[3];
  double* g[3][3];  // g[field][component]
};
void launch_lsq_gradient3(const MeshView& m, const Grad3Args& a, hipStream_t s);

// ----------------------------------------------------------- assembly
void launch_elliptic_assemble(const MeshView& m, const double* gammaFace,
                              const double* cellSource, BCView bc, const double* gx,
                              const double* gy, const double* gz, int includeNonOrth,
                              int useRef, int refCell, double refValue, double* vals,
                              double* rhs, hipStream_t s);
void launch_transport_assemble(const MeshView& m, const double* faceFlux, const double* gammaFace,
                               const double* Su, const double* Sp, BCView bc, const double* gx,
                               const double* gy, const double* gz, int includeNonOrth,
                               int upwind, double* vals, double* rhs, hipStream_t s);

// interleaved <-> SoA helpers
void launch_aos_to_soa(int n, const double* aos, double* x, double* y, double* z, hipStream_t s);
void launch_soa_to_aos(int n, const double* x, const double* y, const double* z, double* aos,
                       hipStream_t s);

}  // namespace an
