/*
 * oracle.h -- CPU FP64 restatement of the Anabasis SIMPLE hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or as the timed CPU baseline).  The product library (anabasis_amd) never
 * links, loads or calls it.
 *
 * Every function restates a reference routine; the file:line it follows is
 * cited next to it (paths relative to the Anabasis source tree).  The
 * arithmetic is written in the reference's own operation order and the
 * library is built with -ffp-contract=off, so that a kernel which evaluates
 * the same expressions in the same order agrees bit for bit.
 *
 * Array conventions: 3-vectors are stored interleaved ([3*i+0..2]); boundary
 * face arrays (bc types / values) are indexed by bf = f - nInternalFaces;
 * per-patch arrays by the 0-based patch index; bPatch[f] holds patch+1 (0 for
 * internal faces) exactly like the reference Mesh::bPatch.
 *
 * Pinning status of each function is listed in the header comment of the .c
 * file that defines it.
 */
#ifndef ANABASIS_ORACLE_H
#define ANABASIS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_BC_NEUMANN 0
#define OR_BC_DIRICHLET 1
#define OR_BC_ZEROGRAD 2

/* ---------------------------------------------------------------- mesh */
/* LP/src/mesh.cu:188-285 (cell centroid/volume, face centre/area/normal,
 * normal orientation, maxNonOrthDeg).  Returns 0, or -(1+c) for a
 * non-positive volume at cell c, or -(1+nCells+f) for a degenerate face f. */
int or_mesh_geometry(int nPoints, const double *P, int nFaces,
                     const int *faceOff, const int *faceVerts,
                     const int *owner, int nInternalFaces, const int *neigh,
                     int nCells, double *cc, double *vol, double *xf,
                     double *nf, double *Sf, double *Af,
                     double *maxNonOrthDeg);

/* LP/src/scalar_elliptic.cu:16-58 (== GSV1 main.cu:1315-1340).
 * nnz is returned by or_pattern_nnz; or_pattern_fill fills caller arrays. */
int64_t or_pattern_nnz(int nCells, int nFaces, int nInternalFaces,
                       const int *owner, const int *neigh);
int or_pattern_fill(int nCells, int nFaces, int nInternalFaces,
                    const int *owner, const int *neigh, int *rowOffsets,
                    int *cols, int *diagPos, int *facePP, int *facePN,
                    int *faceNP, int *faceNN);

/* ---------------------------------------------------------- gradients */
/* GSV1: build_lsq_gradient_operator (main.cu:1159-1251) + apply
 * (kernel_apply_lsq_gradient, main.cu:1403-1440).  bcTypeBF/bcValBF per
 * boundary face; non-Dirichlet boundary terms use phi_P.  grad out [3*nC]. */
void or_lsq_gradient_gsv1(int nCells, int nFaces, int nInternalFaces,
                          const int *owner, const int *rowOffsets,
                          const int *cols, const double *cc, const double *xf,
                          const int *bcTypeBF, const double *bcValBF,
                          const double *phi, double *grad);
/* LP: compute_lsq_gradient (gradient.cu:21-83), identical in
 * LS/src/scalar_transport_library.cu:185-248: Neumann faces use r=dn*n. */
void or_lsq_gradient_lp(int nCells, int nFaces, int nInternalFaces,
                        const int *owner, const int *rowOffsets,
                        const int *cols, const double *cc, const double *xf,
                        const double *nf, const int *bcTypeBF,
                        const double *bcValBF, const double *phi,
                        double *grad);

/* ----------------------------------------------------------- assembly */
/* LP: assemble_scalar_elliptic_system (scalar_elliptic.cu:60-136). */
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
                         double *values, double *rhs);

/* LS: assemble_scalar_transport_system (scalar_transport_library.cu:250-357).
 * upwind != 0 selects ConvectionScheme::Upwind. */
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
                          double *values, double *rhs);

/* GSV1 momentum: kernel_zero_values + kernel_momentum_base_steady +
 * kernel_momentum_internal_faces + kernel_momentum_boundary_faces
 * (main.cu:1673-1684, 1804-1896), accumulated serially in face order.
 * gradQ is [3*nC] interleaved; bcQ/U/V/W types per PATCH, values per FACE
 * (nFaces, as the reference DeviceBC).  Output vals[nnz], rhs[nC]. */
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
                          double *rhs);
/* kernel_momentum_base_rhs_only + *_rhs_only face kernels
 * (main.cu:1697-1785). */
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
                          double *rhs);
/* kernel_relax_momentum_system (main.cu:1686-1695). */
void or_momentum_relax(int nCells, const int *diagPos, double *vals,
                       double *rhs, const double *qOld, double uRelax);
/* kernel_relax_momentum_rhs_only_from_relaxed_matrix (main.cu:1787-1794). */
void or_momentum_relax_rhs_only(int nCells, const int *diagPos,
                                const double *vals, double *rhs,
                                const double *qOld, double uRelax);
/* kernel_extract_vol_over_diag (main.cu:1796-1802). */
void or_extract_rau(int nCells, const int *diagPos, const double *vals,
                    const double *vol, double *rAU);
/* update_pressure_matrix_from_rAU (main.cu:2532-2551 -> 1898-1950). */
void or_pressure_assemble(int nCells, int nFaces, int nInternalFaces,
                          const int *owner, const int *neigh,
                          const int *bPatch, const int *diagPos,
                          const int *facePP, const int *facePN,
                          const int *faceNP, const int *faceNN,
                          const double *cc, const double *xf,
                          const double *nf, const double *Af,
                          const double *rAU, const int *bcPType,
                          int64_t nnz, double *vals);

/* --------------------------------------------------------- SIMPLE ops */
/* kernel_build_rhiechow_predicted_flux_stokes_3d (main.cu:1474-1521). */
void or_rhiechow_flux(int nFaces, int nInternalFaces, const int *owner,
                      const int *neigh, const int *bPatch, const double *cc,
                      const double *xf, const double *nf, const double *Af,
                      const double *u, const double *v, const double *w,
                      const double *p, const double *gradP,
                      const double *rAU, const int *bcUType,
                      const double *uFaceBC, const int *bcVType,
                      const double *vFaceBC, const int *bcWType,
                      const double *wFaceBC, double rho, double *phiStar);
/* kernel_continuity_residual_from_flux (main.cu:1523-1536), face order. */
void or_continuity(int nCells, int nFaces, int nInternalFaces,
                   const int *owner, const int *neigh, const double *phi,
                   double *div);
/* kernel_pressure_nonorth_flux_and_divergence (main.cu:1538-1593). */
void or_pressure_nonorth(int nCells, int nFaces, int nInternalFaces,
                         const int *owner, const int *neigh,
                         const int *bPatch, const double *cc,
                         const double *xf, const double *nf,
                         const double *Sf, const double *Af,
                         const double *rAU, const int *bcPType,
                         const double *gradP, double *phiNonOrth,
                         double *divNonOrth);
/* kernel_correct_face_fluxes_simple_nonorth (main.cu:1615-1658). */
void or_correct_flux(int nFaces, int nInternalFaces, const int *owner,
                     const int *neigh, const int *bPatch, const double *cc,
                     const double *xf, const double *nf, const double *Af,
                     const double *rAU, const int *bcPType,
                     const double *pFaceBC, const double *phiStar,
                     const double *pCorr, const double *phiNonOrth,
                     double *phi);
/* compute_patch_forces_wall_shear (main.cu:2828-2933).  out[0..8] = Fp,Fv,F;
 * out[9] area, out[10] nFaces, out[11] minDn, out[12] maxDn, out[13] maxUt,
 * out[14] maxShear. */
void or_patch_forces(const int *owner, const double *cc, const double *xf,
                     const double *nf, const double *Af, int f0, int f1,
                     const double *u, const double *v, const double *w,
                     const double *p, double mu, int normalSign,
                     double *out);

/* -------------------------------------------------------------- solvers */
void or_spmv(int n, const int *rowOffsets, const int *cols, const double *vals,
             const double *x, double *y);

/* Aggregation-AMG preconditioned CG (the replacement for HYPRE PCG +
 * BoomerAMG; the GPU solver runs the same algorithm).  Stop when
 * ||r||_2 <= max(relTol*||b||_2, absTol) (HYPRE two-norm PCG criterion,
 * LP/src/hypre_backend.cu:55-62).  x holds x0 on entry.  hist (may be NULL)
 * receives ||r||_2 per iteration (index 0 = initial), up to histCap values.
 * Returns iterations (negative if not converged). */
typedef struct {
  int maxIter;
  double relTol;
  double absTol;
  int coarseSize;      /* stop coarsening below this many rows */
  int maxLevels;
  double strength;     /* pairwise matching strength threshold */
  int pairPasses;      /* pairwise passes per level (2 -> ~4:1) */
  double jacobiOmega;  /* Jacobi smoother damping */
  int sweeps;          /* pre = post sweeps */
  int kcycle;          /* coarse levels 1..kcycle (never the coarsest) take
                          their correction from two flexible-CG steps
                          preconditioned by the next cycle (K-cycle); the
                          outer PCG then uses the flexible (Polak-Ribiere)
                          beta.  0: V-cycle */
  int smoother;        /* 0: damped Jacobi (jacobiOmega, sweeps); 1: degree-2
                          Chebyshev on levels 0..chebLevels-1 (all smoothed
                          levels when chebLevels <= 0), Jacobi below */
  int chebLevels;
  double chebFraction; /* Chebyshev interval [chebFraction lmax, lmax] with
                          lmax the Gershgorin bound of D^-1 A */
} or_amg_options;

int or_pcg_amg(int n, const int *rowOffsets, const int *cols,
               const double *vals, const double *b, double *x,
               const or_amg_options *opt, double *finalRelRes,
               double *hist, int histCap);
/* AMG hierarchy introspection for parity tests: fills aggregate ids of level
 * 0 -> 1 (agg[n]) and returns the number of coarse rows. */
int or_amg_aggregate(int n, const int *rowOffsets, const int *cols,
                     const double *vals, const or_amg_options *opt,
                     int *agg);

/* One cycle z = M^-1 r with a freshly built hierarchy; returns levels. */
int or_amg_vcycle(int n, const int *rowOffsets, const int *cols,
                  const double *vals, const or_amg_options *opt,
                  const double *r, double *z);
/* The K-cycle step (two flexible-CG steps preconditioned by the plain cycle)
 * on the given matrix as its own level; returns levels. */
int or_amg_kstep(int n, const int *rowOffsets, const int *cols,
                 const double *vals, const or_amg_options *opt,
                 const double *b, double *x);

/* Reusable PCG + AMG: hierarchy kept between solves; set_values with
 * rebuild == 0 refreshes only the level-0 smoother (stale coarse levels). */
void *or_pcg_create(int n, const int *rowOffsets, const int *cols,
                    const or_amg_options *opt);
void or_pcg_set_values(void *h, const double *vals, int rebuild);
int or_pcg_solve(void *h, const double *b, double *x, double relTol,
                 double absTol, int maxIter, double *finalRelRes);
void or_pcg_destroy(void *h);

/* Jacobi (DiagScale) preconditioned BiCGSTAB, HYPRE stop rule
 * ||r|| <= max(relTol*||b||, absTol).  Returns iterations (negative if not
 * converged). */
int or_bicgstab_jacobi(int n, const int *rowOffsets, const int *cols,
                       const double *vals, const double *b, double *x,
                       int maxIter, double relTol, double absTol,
                       double *finalRelRes);

/* threads the oracle's parallel loops use */
int or_num_threads(void);

#ifdef __cplusplus
}
#endif

#endif
