// amg.hpp -- aggregation AMG hierarchy (replacement for HYPRE BoomerAMG).
#pragma once

#include "linalg.hpp"

namespace an {

struct AmgOptions {
  int coarseSize = 256;
  int maxLevels = 25;
  double strength = 0.25;
  int pairPasses = 2;
  double omega = 0.75;
  int sweeps = 1;  // Jacobi sweeps per side of the coarse correction
  int kcycle = 0;  // levels 1..kcycle (not the coarsest): K-cycle correction
  int smoother = 0;            // 1: degree-2 Chebyshev on the chebLevels finest levels
  int chebLevels = 1;          // <= 0: every smoothed level
  double chebFraction = 0.3;   // lmin = chebFraction * lmax
};

struct AmgLevel {
  int n = 0, nc = 0;
  DevCSR<double> own;           // owned matrix (levels >= 1)
  const DevCSR<double>* A = nullptr;
  DevBuf<double> omdinv;        // omega / a_ii
  DevBuf<int> agg, moff, mem;   // aggregation to level+1
  DevBuf<int> mpos;             // mpos[mem[k]] = k (the down sweep's store order)
  RowBlocks mrb;                // row blocks of (moff, mem) for the restriction
  // degree-2 Chebyshev smoother: chebCoef (device) = {1/theta, cv, cw, cr};
  // omdinv then holds (1/theta) / a_ii
  bool cheb = false;
  DevBuf<double> chebCoef;
  DevBuf<double> coarseInvT;    // coarsest: dense inverse, column-major
  DevBuf<double> b, x, r, y;    // work vectors (FP64); y = omdinv o b
  // FP32 copies for the mixed-precision V-cycle
  DevCSR<float> A32;
  DevBuf<float> omdinv32, b32, x32, r32, y32;
  // K-cycle level: first coarse correction c1 and its product A c1
  DevBuf<double> kc, kv;
  DevBuf<float> kc32, kv32;
  bool kLevel = false;
  bool isCoarsest = false;
};

// Per-precision views of a level's matrix and work vectors.
template <typename T>
struct LevelData;
template <>
struct LevelData<double> {
  static const DevCSR<double>& A(const AmgLevel& L) { return *L.A; }
  static const double* omd(const AmgLevel& L) { return L.omdinv.p; }
  static double* b(AmgLevel& L) { return L.b.p; }
  static double* x(AmgLevel& L) { return L.x.p; }
  static double* r(AmgLevel& L) { return L.r.p; }
  static double* y(AmgLevel& L) { return L.y.p; }
  static double* kc(AmgLevel& L) { return L.kc.p; }
  static double* kv(AmgLevel& L) { return L.kv.p; }
};
template <>
struct LevelData<float> {
  static const DevCSR<float>& A(const AmgLevel& L) { return L.A32; }
  static const float* omd(const AmgLevel& L) { return L.omdinv32.p; }
  static float* b(AmgLevel& L) { return L.b32.p; }
  static float* x(AmgLevel& L) { return L.x32.p; }
  static float* r(AmgLevel& L) { return L.r32.p; }
  static float* y(AmgLevel& L) { return L.y32.p; }
  static float* kc(AmgLevel& L) { return L.kc32.p; }
  static float* kv(AmgLevel& L) { return L.kv32.p; }
};

// Inner products finished by a preconditioner application z = M^-1 r:
// out[0] = r.z and, when q is given (same precision as r), out[1] = z.q.
struct ApplyDots {
  const void* q;
  RedArgs ra;
  bool yReady = false;  // level 0's y = omdinv o r already written (by the PCG update)
};

class AmgHierarchy {
 public:
  // block cache for every buffer the setup allocates (declared first: it
  // outlives the levels whose buffers come from it); rebuilds then reuse
  // blocks instead of paying hipMalloc/hipFree per temporary
  DevPool pool;
  AmgOptions opt;
  std::vector<AmgLevel> lv;
  bool built = false;
  bool has32 = false;

  // Builds the hierarchy from the fine matrix (kept by reference).
  void setup(const DevCSR<double>& fine, hipStream_t s);
  // Refreshes level-0 smoother data after the fine values changed (stale
  // coarse levels, as a HYPRE hierarchy reused between setups).
  void refresh_fine(hipStream_t s);
  void cheb_setup(AmgLevel& L, hipStream_t s);
  void make_fp32(hipStream_t s);
  // z = M^-1 r, one cycle from zero (V, or K on the kcycle levels)
  void apply(const double* r, double* z, hipStream_t s, const ApplyDots* dots = nullptr);
  void apply32(const float* r, float* z, hipStream_t s, const ApplyDots* dots = nullptr);
  int levels() const { return (int)lv.size(); }
  // some level takes the K-cycle correction: the preconditioner is not a
  // fixed linear operator, so the outer CG uses the flexible beta
  bool flexible() const;

  Reducer kred;          // K-cycle inner products (one level at a time)
  DevBuf<double> kscal;  // 8 scalars per level: rho1, c1.b, c2.v1, c2.Ac2, c2.b
};

// Dense inverse of the coarsest level (n <= 4096), built on the GPU:
// Cholesky with pivots below 1e-13*max|diag| treated as zero, then one
// forward and one backward substitution per column of the identity.  Every
// entry is formed by the same sequence of operations as the oracle's
// dense_inverse (oracle/solvers.c), so the result is the same bits.
// invT[j*n + i] = inverse[i][j] (column-major).
void dense_inverse_device(const DevCSR<double>& A, DevBuf<double>& invT, hipStream_t s);

}  // namespace an
