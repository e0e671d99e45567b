// simple.hpp -- device-resident SIMPLE outer iteration (generic_simple_v1).
#pragma once

#include <memory>
#include <string>

#include "kernels.hpp"
#include "solver.hpp"

namespace an {

// inputs of the momentum assembly kernel (one velocity component)
struct MomArgs {
  const int* bcT;          // velocity BC type per boundary face
  const double* qB;        // component BC value per boundary face
  const double *uB, *vB, *wB;
  const double *gqx, *gqy, *gqz;  // gradient of the current iterate
  const double *uC, *vC, *wC;     // convecting velocity (old fields)
  const double* gpc;       // pressure-gradient component
  const double* qOld;      // iterate used by the relaxation
  double rho, mu, corrPsi, uRelax;
};

class SimpleSolver {
 public:
  SimpleSolver(DevMesh& mesh, const an_simple_params& par, const int* patchUType,
               const int* patchPType, const double* uFaceBC, const double* vFaceBC,
               const double* wFaceBC, const double* pFaceBC, int inletPatch);
  ~SimpleSolver();
  // Runs n outer iterations; returns AN_EDIVERGED on divergence.
  an_status iterate(int n, an_simple_stats* last, int* converged);
  void get_field(const std::string& name, double* out, bool device);
  void set_field(const std::string& name, const double* in, bool device);
  void patch_forces(int patch, int normalSign, double* out);
  void timings(double* out);

  DevMesh& M;
  MeshView mv;
  an_simple_params par;
  hipStream_t s = 0;
  int step = 0;
  bool converged = false;

 private:
  // boundary data per boundary face
  DevBuf<int> bcU, bcP;
  DevBuf<double> bcUv[3], bcPv;
  // fields
  DevBuf<double> q[3], qOld[3], qNew, p, pOld, pCorr, rAU;
  DevBuf<double> gp[3], g[3];
  DevBuf<double> phiStar, phi, phiNo, divStar, divCorr, rhs;
  DevCSR<double> Amom;
  DevBuf<double> pvals;
  DevBuf<double> pCoef;  // per internal face: A rAU_f / dpn (0 where skipped)
  std::unique_ptr<PcgSolver> psolver;
  BicgWork bw;
  // fused u/v/w momentum (par.fuseMomentumRhs): per-component gradients,
  // right-hand sides and new iterates
  Bicg3Work bw3;
  DevBuf<double> g3[3][3], rhs3[3], qNew3[3];
  DevBuf<double> momCorr, momCoef;  // per internal face: 3 corrections, 4 matrix terms
  DevBuf<double> momDinv;           // 1 / diagonal of the momentum matrix (Jacobi)
  Reducer red;
  DevBuf<double> scal;
  PinnedScalars hs;
  double tAcc[6] = {0, 0, 0, 0, 0, 0};
  long long pcgCum = 0, pSolveCum = 0;  // pressure PCG iterations / solves
  hipEvent_t ev[8];

  void grad(int bcKind, const double* f, DevBuf<double>* out);
  void momentum_component(int comp, an_simple_stats* st);
  void momentum_fused(an_simple_stats* st);
  MomArgs mom_args(int comp, const DevBuf<double>* grad) const;
  double relchg(const double* a, const double* b);  // device reduction
  // several reductions with one copy and one synchronisation: queue_stats
  // reduces a (and a - b) into slot k (k < kStatSlots); fetch_stats brings
  // slots 0..nk-1 to the host, read with stat_relchg / stat_maxabs
  static constexpr int kStatSlots = 5;
  void queue_stats(const double* a, const double* b, int k);
  void fetch_stats(int nk);
  double stat_relchg(int k) const;
  double stat_maxabs(int k) const;
  void mark(int i) { AN_HIP(hipEventRecord(ev[i], s)); }
  double since(int a, int b);
};

}  // namespace an
