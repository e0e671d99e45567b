/*
 * anabasis_hip.h -- C ABI of the MI355X-native Anabasis SIMPLE hot path.
 *
 * Plain C: opaque handles, int32 indices, double/float values, explicit
 * lengths.  Every entry point returns an an_status (AN_OK == 0); the message
 * of the last failure on the calling thread is an_last_error().  No C++
 * exception crosses this boundary.  Non-convergence of an iterative solve is
 * reported (AN_ENOCONV) and the approximate solution is still returned, as
 * the reference does for HYPRE error 256 (GSV1 main.cu:2656-2665).
 *
 * Array conventions (the reference's own, flattened):
 *   - 3-vectors are interleaved [3*i+0..2] (std::array<double,3> layout of
 *     LP Mesh::cc etc.) unless the name says SoA;
 *   - "BF" arrays are per boundary face, index bf = f - nInternalFaces
 *     (LP BoundaryFaceData holds the same entries at index f);
 *   - patch indices are 0-based; bPatch[f] is patch+1 (0 = internal face),
 *     exactly LP Mesh::bPatch (LP/src/mesh.cu:179-186).
 * Pointers marked "host" are host memory; functions with an an_memspace
 * argument accept either host or device pointers for their bulk arrays.
 * Reference files are cited relative to the Anabasis source tree.
 */
#ifndef ANABASIS_HIP_H
#define ANABASIS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  AN_OK = 0,
  AN_EINVAL = 1,    /* bad argument / mesh / size mismatch */
  AN_ENOCONV = 2,   /* iterative solve did not reach its tolerance */
  AN_EHIP = 3,      /* HIP runtime error */
  AN_EIO = 4,       /* file could not be read or written */
  AN_ENOMEM = 5,
  AN_EDIVERGED = 6, /* SIMPLE mass residual non-finite or > 1e50 */
  AN_EINTERNAL = 7
} an_status;

typedef enum { AN_HOST = 0, AN_DEVICE = 1 } an_memspace;

/* boundary condition kinds (LP ScalarBCType, LS ScalarBCType, GSV1
 * bc_to_int main.cu:1058) */
#define AN_BC_NEUMANN 0
#define AN_BC_DIRICHLET 1
#define AN_BC_ZERO_GRADIENT 2

const char *an_last_error(void);
int an_version(void);
/* Selects the HIP device used by later calls on this thread. */
an_status an_set_device(int device);

/* ------------------------------------------------------------------ mesh */
typedef struct an_mesh_s *an_mesh_t;

typedef struct {
  int nPoints, nFaces, nInternalFaces, nCells, nPatches;
  int nBoundaryFaces;
  int64_t nnz;          /* CSR nonzeros of the cell-to-cell pattern */
  double maxNonOrthDeg; /* LP Mesh::maxNonOrthDeg */
  int hasGeometry;
} an_mesh_info;

/* read_openfoam_polymesh (LP/src/mesh.cu:160-288; GSV1 main.cu:582-599).
 * After parsing, the mesh and its geometry are also written to
 * <polyMeshDir>/.anabasis_mesh.bin (when the directory is writable); later
 * reads load that file instead while the five ASCII files keep the size,
 * modification time and content fingerprint it records (2.08M cells: 6.5 s
 * -> 0.7 s).  an_mesh_write_polymesh removes it; ANABASIS_MESH_CACHE=0 in
 * the environment switches the cache off.  No reference counterpart. */
an_status an_mesh_read_polymesh(const char *polyMeshDir, an_mesh_t *out);
/* Mesh from arrays (host): faces as offsets[nFaces+1] + vertex ids.
 * nPoints == 0 / points == NULL gives a topology-only mesh (CSR addressing
 * and patch lists only; no geometry). */
an_status an_mesh_from_arrays(int nPoints, const double *points, int nFaces,
                              const int *faceOffsets, const int *faceVerts,
                              const int *owner, int nInternalFaces,
                              const int *neigh, int nPatches,
                              const char *const *patchNames,
                              const int *patchStart, const int *patchNFaces,
                              an_mesh_t *out);
an_status an_mesh_write_polymesh(an_mesh_t mesh, const char *dir);
/* Generators (no reference counterpart: the reference ships no generator).
 * hex box: nx*ny*nz hexes on [0,lx]x[0,ly]x[0,lz], interior vertices
 * jittered uniformly by +-jitter*h (boundary vertices stay on their faces),
 * mt19937_64(seed); patches xmin,xmax,ymin,ymax,zmin,zmax. */
an_status an_mesh_gen_hexbox(int nx, int ny, int nz, double lx, double ly,
                             double lz, double jitter, uint64_t seed,
                             an_mesh_t *out);
/* unit-cube tet mesh: n^3 blocks, 6 Kuhn tets each; same patch names. */
an_status an_mesh_gen_tetcube(int nblocks, double length, an_mesh_t *out);
/* 2.2 x 0.41 x 0.41 channel with a D=0.1 cylinder (axis z, centre
 * (0.5,0.2)), O-grid + H blocks, about targetCells hexes; patches in the
 * order of examples/cylinder: patch_1_0 inlet, patch_0_0 walls y=0/H,
 * patch_2_0 outlet, patch_5_0 walls z=0/W, patch_3_0 cylinder. */
an_status an_mesh_gen_cylinder(long long targetCells, an_mesh_t *out);
an_status an_mesh_destroy(an_mesh_t mesh);
an_status an_mesh_get_info(an_mesh_t mesh, an_mesh_info *info);
/* Copies a host mesh array.  name: "points","faceOffsets","faceVerts",
 * "owner","neigh","bPatch","cc","vol","xf","nf","Sf","Af".  out must hold
 * the documented number of elements (int or double). */
an_status an_mesh_get_array(an_mesh_t mesh, const char *name, void *out);
/* Patch k: name (nul-terminated, truncated to cap), startFace, nFaces. */
an_status an_mesh_get_patch(an_mesh_t mesh, int k, char *name, int cap,
                            int *startFace, int *nFaces);

/* ------------------------------------------------------------ addressing */
/* build_scalar_pattern (LP/src/scalar_elliptic.cu:16-58, == GSV1
 * build_momentum_pattern main.cu:1315-1340), built on the GPU.  Outputs
 * rowOffsets[nCells+1], cols[nnz], diagPos[nCells], facePP/PN/NP/NN[nIF]
 * (any may be NULL). */
an_status an_mesh_csr(an_mesh_t mesh, int *rowOffsets, int *cols,
                      int *diagPos, int *facePP, int *facePN, int *faceNP,
                      int *faceNN, an_memspace mem);
/* Same pattern from bare owner/neighbour lists (e.g. a polyMesh without
 * points/faces).  nnzOut receives nnz; when the output pointers are NULL
 * only nnz is computed. */
an_status an_csr_from_faces(int nCells, int nFaces, int nInternalFaces,
                            const int *owner, const int *neigh,
                            int64_t *nnzOut, int *rowOffsets, int *cols,
                            int *diagPos, int *facePP, int *facePN,
                            int *faceNP, int *faceNN, an_memspace mem);

/* ------------------------------------------------------------- gradients */
#define AN_GRAD_GSV1 0 /* kernel_apply_lsq_gradient, GSV1 main.cu:1403-1440 */
#define AN_GRAD_LP 1   /* compute_lsq_gradient, LP/src/gradient.cu:21-83 */
/* Least-squares cell gradient of phi[nCells] -> grad[3*nCells] interleaved.
 * bcTypeBF/bcValBF: per boundary face. */
an_status an_lsq_gradient(an_mesh_t mesh, int mode, const int *bcTypeBF,
                          const double *bcValBF, const double *phi,
                          double *grad, an_memspace mem);

/* -------------------------------------------------------------- assembly */
/* assemble_scalar_elliptic_system (LP/src/scalar_elliptic.cu:60-136):
 * gammaFace[nFaces], cellSource[nCells], grad[3*nCells] -> values[nnz]
 * (CSR order of an_mesh_csr), rhs[nCells]. */
an_status an_elliptic_assemble(an_mesh_t mesh, const double *gammaFace,
                               const double *cellSource, const int *bcTypeBF,
                               const double *bcValBF, const double *grad,
                               int includeNonOrth, int useReferenceCell,
                               int refCell, double refValue, double *values,
                               double *rhs, an_memspace mem);
/* assemble_scalar_transport_system (LS/src/scalar_transport_library.cu:
 * 250-357); upwind != 0 selects ConvectionScheme::Upwind. */
an_status an_transport_assemble(an_mesh_t mesh, const double *faceFlux,
                                const double *gammaFace, const double *Su,
                                const double *Sp, const int *bcTypeBF,
                                const double *bcValBF, const double *grad,
                                int includeNonOrth, int upwind,
                                double *values, double *rhs, an_memspace mem);

/* ---------------------------------------------------------- linear solve */
typedef struct {
  int maxIter;         /* HypreOptions::maxIter */
  double relTol;       /* stop when ||r|| <= max(relTol*||b||, absTol) */
  double absTol;
  int precision;       /* 0 FP64; 1 FP32 inner PCG+AMG in FP64 refinement;
                        * 2 FP64 PCG (SpMV, residuals, inner products, stop
                        * test) preconditioned by the FP32 AMG cycle */
  int coarseSize;      /* AMG: coarsest level size (dense solve) */
  int maxLevels;
  double strength;     /* AMG: pairwise matching strength threshold */
  int pairPasses;      /* AMG: matchings per level */
  double jacobiOmega;  /* AMG: damped-Jacobi smoother weight */
  double innerRelTol;  /* mixed precision: inner FP32 solve reduction */
  int profile;         /* 1: time every SpMV / V-cycle with HIP events */
  int sweeps;          /* AMG: damped-Jacobi sweeps before and after the
                          coarse correction on every level (>= 1) */
  int kcycle;          /* AMG: coarse levels 1..kcycle (never the coarsest)
                          take their correction from two flexible-CG steps
                          preconditioned by the next cycle (K-cycle); the
                          outer CG then uses the flexible (Polak-Ribiere)
                          beta.  0: V-cycle.  Ignored by the batched solver */
  int graphs;          /* 1: each preconditioner application is replayed
                          from a captured HIP graph (captured once per
                          hierarchy build) -- for hosts too slow to keep the
                          GPU fed; 0 (default): launched kernel by kernel */
  int rhsProjection;   /* successive right-hand sides with one matrix: keep
                          up to this many (<= 16) previous solutions X
                          (A X orthonormal) and move x0 to the point of
                          x0 + span X with the smallest residual before
                          iterating; the set is cleared when matrix values
                          are set.  0 (default): off */
  int smoother;        /* AMG: 0 damped Jacobi (jacobiOmega, sweeps); 1
                          degree-2 Chebyshev on the chebLevels finest levels
                          (all smoothed levels when chebLevels <= 0), on the
                          interval [chebFraction lmax, lmax] of D^-1 A with
                          lmax its Gershgorin bound; Jacobi below.  The
                          batched solver takes 0 only */
  int chebLevels;
  double chebFraction;
} an_pcg_options;
void an_pcg_default_options(an_pcg_options *o);

typedef struct {
  int iterations;          /* HypreSolveInfo::iterations */
  double finalRelResNorm;  /* HypreSolveInfo::finalRelResNorm */
  double finalResNorm;
  int converged;
  int levels;              /* AMG levels */
  double setupSeconds, solveSeconds;
  int refinementSteps;     /* mixed precision outer steps */
  /* with an_pcg_options.profile: in-loop kernel timings (HIP events on the
   * solver stream) of the Krylov SpMV and of one preconditioner apply */
  int spmvCalls;
  double spmvAvgSeconds;
  int vcycleCalls;
  double vcycleAvgSeconds;
  /* bytes of column index the Krylov SpMV reads per nonzero: 2 when every
   * column of the matrix fits a 16-bit offset from its CSR row block's
   * first row (banded mesh orderings), else 4 */
  int spmvIndexBytes;
} an_solve_info;

typedef struct an_solver_s *an_solver_t;
/* PCG + aggregation-AMG on a CSR pattern (the replacement for HYPRE PCG +
 * BoomerAMG, LP/src/hypre_backend.cu:3-114 and the reusable API
 * LP/include/hypre_backend.h:64-82).  The pattern is copied. */
an_status an_pcg_create(int n, const int *rowOffsets, const int *cols,
                        int64_t nnz, const an_pcg_options *opt,
                        an_memspace mem, an_solver_t *out);
/* New matrix values; rebuildHierarchy != 0 also rebuilds the AMG hierarchy
 * (update_reusable_hypre_system_gpu doSetup). */
an_status an_pcg_set_values(an_solver_t s, const double *values,
                            int rebuildHierarchy, an_memspace mem);
/* Solves A x = b; x holds the initial guess on entry. */
an_status an_pcg_solve(an_solver_t s, const double *b, double *x,
                       an_memspace mem, an_solve_info *info);
/* Runs the handle's later work (an_pcg_set_values, an_pcg_solve, an_spmv,
 * an_amg_apply) on the caller's HIP stream (a hipStream_t; NULL = the null
 * stream, the default).  Device-pointer inputs are then read in that
 * stream's order, so a caller need not synchronise before a call; each call
 * still returns with its results complete, as the reference's synchronous
 * HYPRE calls do (LP/src/hypre_backend.cu:35-114 on the legacy default
 * stream).  Work already queued on the previous stream is finished first.
 * The stream must outlive its use by the handle. */
an_status an_pcg_set_stream(an_solver_t s, void *hipStream);
/* Residual 2-norms of the last an_pcg_solve: ||r0|| and one per PCG
 * iteration (precision 1: ||r0|| and the FP64 residual after each
 * refinement step), the values HYPRE prints at print level 2 when the
 * reference's `monitor` option is set (LP/src/hypre_backend.cu:63-64,
 * 129-130).  Copies min(cap, count) values into out (host memory); *count
 * receives the full length. */
an_status an_pcg_residual_history(an_solver_t s, double *out, int cap, int *count);
/* y = A x with the solver's current values (test hook for the SpMV). */
an_status an_spmv(an_solver_t s, const double *x, double *y, an_memspace mem);
/* One preconditioner application z = M^-1 r (test hook for the V-cycle). */
an_status an_amg_apply(an_solver_t s, const double *r, double *z,
                       an_memspace mem);
/* Level-0 aggregate ids (host int[n]); returns the coarse size. */
an_status an_amg_aggregates(an_solver_t s, int *agg, int *nCoarse);
an_status an_solver_destroy(an_solver_t s);

/* BiCGSTAB + Jacobi (HYPRE BiCGSTAB + DiagScale, GSV1 main.cu:2553-2576,
 * LS/src/scalar_transport_library.cu:359-471), one-shot. */
an_status an_bicgstab_solve(int n, const int *rowOffsets, const int *cols,
                            int64_t nnz, const double *values,
                            const double *b, double *x, int maxIter,
                            double relTol, double absTol, an_memspace mem,
                            an_solve_info *info);

/* an_bicgstab_solve without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under
 * stream capture, tight serving loops, hosts that forbid hipMalloc).
 * rowOffsets[n+1], cols[nnz], values[nnz], b[n] and x[n] (x0 on entry) are
 * DEVICE pointers; the result is the same as an_bicgstab_solve with
 * AN_DEVICE on the same inputs, bit for bit.  Every work buffer comes from
 * one AN_NOALLOC_SCRATCH_BYTES device array per GPU that the library holds
 * until the process ends, so n is limited to AN_NOALLOC_MAX_ROWS (AN_EINVAL
 * beyond it, before any device work).  Calls on one GPU share that array:
 * calls that overlap on the same GPU race.  Runs on the null stream and
 * returns when the solve has finished. */
#define AN_NOALLOC_SCRATCH_BYTES ((size_t)32 << 20)
#define AN_NOALLOC_MAX_ROWS 400000
an_status an_bicgstab_solve_noalloc(int n, const int *rowOffsets, const int *cols,
                                    int64_t nnz, const double *values,
                                    const double *b, double *x, int maxIter,
                                    double relTol, double absTol,
                                    an_solve_info *info);

/* Three right-hand sides sharing one matrix -- the u/v/w momentum systems,
 * which the reference solves one after the other with the x-momentum matrix
 * (GSV1 main.cu:3401-3405, solve_momentum_gpu :2553-2576).  Each column has
 * its own BiCGSTAB scalars and stop test and is bit-identical to
 * an_bicgstab_solve on that column; the fused SpMV reads the matrix once for
 * all three.  b[3], x[3] (x holds x0), info[3]; info[c].spmvAvgSeconds holds
 * the duration of the first fused SpMV.  AN_ENOCONV if any column did not
 * converge. */
an_status an_bicgstab_solve3(int n, const int *rowOffsets, const int *cols,
                             int64_t nnz, const double *values,
                             const double *const *b, double *const *x,
                             int maxIter, double relTol, double absTol,
                             an_memspace mem, an_solve_info *info);
/* ------------------------------------------------ libpoisson / libscalar */
typedef struct {
  int nNonOrthCorr;      /* EllipticOptions::nNonOrthCorr */
  int nonOrth;           /* laplacianScheme: 1 nonorth, 0 orth */
  int useReferenceCell;
  int referenceCell;
  double referenceValue;
  an_pcg_options solver; /* EllipticOptions::hypre */
} an_elliptic_options;
void an_elliptic_default_options(an_elliptic_options *o);
/* solve_scalar_elliptic (LP/src/poisson_library.cu:3-47): non-orthogonal
 * outer loop, each pass from x0 = 0 (hypre_backend.cu:35,49).  Inputs host
 * or device per mem; phi[nCells] out. */
an_status an_solve_scalar_elliptic(an_mesh_t mesh, const double *gammaFace,
                                   const double *cellSource,
                                   const int *bcTypeBF, const double *bcValBF,
                                   const an_elliptic_options *opt,
                                   double *phi, an_memspace mem,
                                   an_solve_info *info);

typedef struct {
  int upwind;          /* ConvectionScheme */
  int nonOrth;         /* DiffusionScheme */
  int nNonOrthCorr;
  int maxIter;
  double relTol, absTol;
} an_transport_options;
void an_transport_default_options(an_transport_options *o);
/* libscalar::solve_steady_scalar_transport
 * (LS/src/scalar_transport_library.cu:475-512); x0 may be NULL. */
an_status an_solve_steady_scalar_transport(
    an_mesh_t mesh, const double *faceFlux, const double *gammaFace,
    const double *Su, const double *Sp, const int *bcTypeBF,
    const double *bcValBF, const an_transport_options *opt, const double *x0,
    double *phi, an_memspace mem, an_solve_info *info);

/* ----------------------------------------------------------------- batch */
/* Many independent sparse SPD systems (e.g. the MMS Poisson problems of
 * apps/poisson_mms_lib run on a set of meshes) solved together: one
 * block-diagonal CSR, one aggregation-AMG hierarchy that never couples
 * problems, and per-problem PCG scalars, stop tests and iteration counts.
 * The PCG iterations run as one persistent launch or as a replayed HIP
 * graph (an_batch_set_launch).  No reference counterpart (the reference
 * solves one system per HYPRE call, LP/src/hypre_backend.cu:3-114).
 * opt: as an_pcg_create (precision 0 or 1; 1 = FP32 PCG+AMG inside FP64
 * refinement).  Patterns are copied. */
typedef struct an_batch_s *an_batch_t;
an_status an_batch_create(int nProblems, const int *nRows,
                          const int *const *rowOffsets, const int *const *cols,
                          const an_pcg_options *opt, an_memspace mem,
                          an_batch_t *out);
/* Values of every problem; (re)builds the hierarchy and runs the
 * autotuner: the block height of the batched kernels (64, 32, 16 or 4 rows
 * per wave, or a wave per row), timed on this matrix, and the launch
 * (an_batch_set_launch). */
an_status an_batch_set_values(an_batch_t b, const double *const *values,
                              an_memspace mem);
/* Solves every problem; x[p] holds x0 on entry; info[nProblems]. */
an_status an_batch_solve(an_batch_t b, const double *const *rhs,
                         double *const *x, an_memspace mem,
                         an_solve_info *info);
/* Autotuner record: chosen rows per wave, and the timed SpMV seconds of the
 * three candidates (64, 32, 16 rows) in seconds[3]. */
an_status an_batch_tuning(an_batch_t b, int *rowsPerWave, double *seconds);
/* enable != 0 (default): the tuner's launch (AN_BATCH_AUTO); 0: issue the
 * launches one by one (AN_BATCH_EAGER). */
an_status an_batch_use_graph(an_batch_t b, int enable);
/* How a batched solve runs its PCG iterations; every mode gives the same
 * bits.  AN_BATCH_PERSISTENT: ONE kernel launch (captured once in a HIP
 * graph) whose co-resident workgroups run every phase of every iteration
 * -- SpMV, per-problem sums and stop tests, each V-cycle level, direction
 * -- separated by grid barriers, until no problem is active.  Its barrier
 * waits are bounded (2 s each): if the workgroups cannot all be resident
 * (another kernel holding the GPU that long), the solve returns AN_EHIP
 * instead of waiting.  AN_BATCH_GRAPH: a captured graph of four iterations
 * replayed until no problem is active (one host read per replay).
 * AN_BATCH_AUTO (default): the faster of the two on this matrix, timed by
 * the tuner in an_batch_set_values. */
#define AN_BATCH_AUTO 0
#define AN_BATCH_EAGER 1
#define AN_BATCH_GRAPH 2
#define AN_BATCH_PERSISTENT 3
an_status an_batch_set_launch(an_batch_t b, int launch);
/* The autotuner's record (an_batch_set_values): block height chosen and
 * the SpMV time of each candidate; launch chosen and the time of one PCG
 * iteration by graph replay and by the persistent kernel at 1, 2 and 4
 * workgroups per CU (0 where that grid cannot be resident). */
typedef struct {
  int rowsPerWave;
  double rowSeconds[5];      /* 64, 32, 16, 4 rows per wave, and one wave per row */
  int launch;                /* AN_BATCH_GRAPH or AN_BATCH_PERSISTENT */
  int persistentBlocksPerCU; /* grid of the persistent kernel */
  double launchSeconds[4];   /* graph, persistent at 1, 2, 4 per CU */
  int launchesLastSolve;     /* kernel-sequence launches of the last solve */
} an_batch_tuning_info;
an_status an_batch_get_tuning(an_batch_t b, an_batch_tuning_info *out);
an_status an_batch_destroy(an_batch_t b);
/* --------------------------------------------------------------- adjoint */
/* Discrete adjoint of libscalar::solve_steady_scalar_transport
 * (LS/src/scalar_transport_library.cu:475-512; assembly :250-357).  No
 * counterpart in the reference, which has no adjoint.
 *
 * For any functional J(phi) of the converged transport solution phi (the
 * non-orthogonal fixed point, or the single orthogonal solve when
 * opt->nonOrth == 0), given dJdphi[nCells] at phi, returns
 * dJdgammaFace[nFaces] = dJ/d gammaFace, dJdbcBF[nBF] = dJ/d bcValBF (the
 * Dirichlet values, or the NeumannGradient values), and optionally the
 * adjoint field[nCells].  The adjoint system A^T psi = dJ/dphi + (dc/dphi)^T
 * psi is iterated opt->nNonOrthCorr+1 times (the transpose of the forward
 * deferred-correction loop), each pass a BiCGSTAB+Jacobi solve with opt's
 * tolerances; the correction's transpose recomputes the least-squares
 * operator instead of storing it. */
an_status an_transport_adjoint(an_mesh_t mesh, const double *faceFlux,
                               const double *gammaFace, const double *Su,
                               const double *Sp, const int *bcTypeBF,
                               const double *bcValBF,
                               const an_transport_options *opt,
                               const double *phi, const double *dJdphi,
                               double *dJdgammaFace, double *dJdbcBF,
                               double *adjoint, an_memspace mem,
                               an_solve_info *info);
/* Per-cell diffusivity -> face diffusivity: gamma_f = (1-lambda_f) k_P +
 * lambda_f k_N on internal faces (lambda_f the face interpolation weight of
 * LS face_interp_lambda, scalar_transport_library.cu:128-135), k_P on
 * boundary faces.  No reference counterpart (the reference takes gammaFace). */
an_status an_face_gamma_from_cells(an_mesh_t mesh, const double *kCell,
                                   double *gammaFace, an_memspace mem);
/* Chain rule of an_face_gamma_from_cells: dJdkCell = (d gamma/d k)^T dJdgammaFace. */
an_status an_dgamma_to_cells(an_mesh_t mesh, const double *dJdgammaFace,
                             double *dJdkCell, an_memspace mem);
/* Transpose of the LP least-squares gradient (an_lsq_gradient mode
 * AN_GRAD_LP, LP/src/gradient.cu:21-83) applied to w[3*nCells]:
 * phibar = (d grad/d phi)^T w, phibarBF = (d grad/d bcValBF)^T w. */
an_status an_lsq_gradient_adjoint(an_mesh_t mesh, const int *bcTypeBF,
                                  const double *bcValBF, const double *w,
                                  double *phibar, double *phibarBF,
                                  an_memspace mem);
/* ------------------------------------------------- runtime boundary conditions */
/* The runtime BC configuration of generic_simple_v1: `velocity <patch>
 * <type> ...` and `pressure <patch> <type> ...` lines, '#' comments.
 * Velocity types wall_noslip, zero_gradient, fixed_uniform_vector ux uy uz,
 * fixed_normal_speed v [mode], fixed_flow_rate Q [mode],
 * parabolic_box_inlet H Umax coordA coordB [mode] (mode
 * average_patch_normal | local_face_normal); pressure types zero_gradient,
 * open, fixed_value v.  Parse errors return AN_EINVAL with the reference's
 * message ("BC config parse error in '<source>' at line <n>: ..."). */
typedef struct an_bc_s *an_bc_t;
/* load_runtime_bc_config (GSV1 bc_runtime_config.cu:83-210). */
an_status an_bc_load(const char *path, an_bc_t *out);
/* The same grammar from a string; sourceName stands for the path in
 * messages. */
an_status an_bc_parse(const char *text, const char *sourceName, an_bc_t *out);
an_status an_bc_destroy(an_bc_t bc);
/* build_patch_geometry_table (GSV1 patch_geometry.cu:83-140) for patch k:
 * out[18] = {nFaces, area, centroid[3], sumSf[3], averageNormal[3],
 * planarity, tangent1[3], tangent2[3]} (serial face-order sums). */
an_status an_patch_geometry(an_mesh_t mesh, int k, double *out);
/* The patch coverage table (GSV1 main.cu:3178-3233).  which = 0: the table,
 * one "  %-32s velocity=yes|NO pressure=yes|NO" line per patch in mesh
 * order; which = 1: the error report printed when a patch lacks a BC (empty
 * when complete).  text receives min(cap-1, length) characters plus a nul;
 * *length the full length; *complete 1 or 0.  Duplicate specs return
 * AN_EINVAL ("Duplicate velocity BC specs found:\n  <patch>\n"). */
an_status an_bc_coverage(an_mesh_t mesh, an_bc_t bc, int which, char *text, int cap,
                         int *length, int *complete);
/* validate_runtime_bc_config_against_patches + duplicate checks +
 * apply_bc_specs_to_legacy_face_arrays (GSV1 bc_runtime_config.cu:212-230,
 * main.cu:3115-3141, velocity_bc_eval.cu:197-290): per-patch types
 * patchUType/patchPType[nPatches] (AN_BC_DIRICHLET or AN_BC_NEUMANN, the
 * reference's bc_to_int) and per-face values [nFaces] (host arrays), the
 * inputs of an_simple_create. */
an_status an_bc_apply(an_mesh_t mesh, an_bc_t bc, int *patchUType, int *patchPType,
                      double *uFaceBC, double *vFaceBC, double *wFaceBC, double *pFaceBC);

/* ------------------------------------------------------------------ SIMPLE */
typedef struct {
  double rho, mu;
  int nVelNonOrthCorr, nNonOrthCorr;
  double uRelax, pRelax;
  double corrTol, tolMass, tolVel;
  int velMaxit;
  double velTol, velRelTol;
  int pMaxit;
  double pTol, pRelTol;
  int pAmgRebuildEvery;
  int precision;        /* pressure solve: as an_pcg_options.precision */
  int fuseMomentumRhs;  /* 1: u/v/w share one 3-RHS BiCGSTAB per pass */
  an_pcg_options pressure;
} an_simple_params;
void an_simple_default_params(an_simple_params *p);

typedef struct {
  int step;
  double massRes, duRel, dvRel, dwRel, dpRel;
  int bicgstabIts[3];
  int pcgLast, pcgTotal;
  double seconds;
} an_simple_stats;

typedef struct an_simple_s *an_simple_t;
/* GSV1 driver state (main.cu:3294-3347).  Per-PATCH velocity/pressure BC
 * types (AN_BC_*), per-FACE BC values [nFaces] (the reference's DeviceBC).
 * u,v,w start at the arithmetic mean of the inlet faces' values when
 * inletPatch >= 0 (main.cu:3302-3322), else 0; p = 0. */
an_status an_simple_create(an_mesh_t mesh, const an_simple_params *par,
                           const int *patchUType, const int *patchPType,
                           const double *uFaceBC, const double *vFaceBC,
                           const double *wFaceBC, const double *pFaceBC,
                           int inletPatch, an_simple_t *out);
/* Runs n outer iterations of main.cu:3388-3634; last-iteration stats.
 * Returns AN_EDIVERGED on divergence, AN_OK otherwise (converged flag in
 * *converged when not NULL). */
an_status an_simple_iterate(an_simple_t s, int n, an_simple_stats *last,
                            int *converged);
/* name: "u","v","w","p","pCorr","rAU","phi"(nFaces),"div", and "pA": the
 * last pressure matrix's values in the CSR order of an_csr_from_faces
 * (nCells + 2 x the number of distinct neighbour pairs) */
an_status an_simple_get_field(an_simple_t s, const char *name, double *out,
                              an_memspace mem);
an_status an_simple_set_field(an_simple_t s, const char *name,
                              const double *in, an_memspace mem);
/* compute_patch_forces_wall_shear (GSV1 main.cu:2828-2933) on the current
 * fields, reduced on the GPU in FP64.  out[15]: Fp[3], Fv[3], F[3], area,
 * nFaces, minWallDistance, maxWallDistance, maxUt, maxShearMag. */
an_status an_simple_patch_forces(an_simple_t s, int patch, int normalSign,
                                 double *out);
/* Accumulated seconds per phase (hipEvent timed): out[0] momentum assembly,
 * [1] momentum solve, [2] pressure assembly+setup, [3] pressure solve,
 * [4] other kernels, [5] total; then counts over the same iterations:
 * [6] pressure PCG iterations, [7] pressure solves (out holds 8 values). */
an_status an_simple_timings(an_simple_t s, double *out);
an_status an_simple_destroy(an_simple_t s);

#ifdef __cplusplus
}
#endif
#endif
