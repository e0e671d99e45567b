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
};
MeshView view_of(const DevMesh& m);

// Per-boundary-face BC (type AN_BC_*, value), device arrays of nBF.
struct BCView {
  const int* type;
  const double* val;
};

// ---------------------------------------------------------- gradients
// mode 0: GSV1 operator semantics, 1: LP compute_lsq_gradient.  Output SoA.
void launch_lsq_gradient(const MeshView& m, int mode, BCView bc, const double* phi, double* gx,
                         double* gy, double* gz, hipStream_t s);

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
