// solver.hpp -- PCG + AMG (FP64 and FP32-inner/FP64-refined) and BiCGSTAB.
#pragma once

#include <map>
#include <memory>
#include <utility>

#include "../../include/anabasis_hip.h"
#include "amg.hpp"

namespace an {

struct PinnedScalars {
  double* h = nullptr;
  double* d = nullptr;  // the same memory as kernels address it (pinned only)
  // pinned == false: h is a plain host array (no allocation at all; copies
  // into it are staged by the runtime)
  explicit PinnedScalars(bool pinned = true) {
    if (pinned) {
      AN_HIP(hipHostMalloc((void**)&h, sizeof(double) * S_COUNT));
      AN_HIP(hipHostGetDevicePointer((void**)&d, h, 0));
      owned_ = true;
    } else {
      h = local_;
    }
    for (int i = 0; i < S_COUNT; ++i) h[i] = 0.0;
  }
  ~PinnedScalars() {
    if (owned_) (void)hipHostFree(h);
  }
  PinnedScalars(const PinnedScalars&) = delete;
  PinnedScalars& operator=(const PinnedScalars&) = delete;

 private:
  bool owned_ = false;
  double local_[S_COUNT];
};

// One 64-bit word in coherent host memory that a kernel stores with a
// system-scope release (pcg_word layout: stamp, decision, float hint); the
// host reads it with acquire loads until the stamp it waits for shows up,
// instead of a copy and a stream synchronisation per iteration.
struct PubWord {
  unsigned long long* h = nullptr;  // host view
  unsigned long long* d = nullptr;  // device view of the same memory
  // kValues doubles after the word (same allocation) for values that a
  // kernel publishes with the word
  static constexpr int kValues = 15;
  double* vh = nullptr;
  double* vd = nullptr;
  PubWord();
  ~PubWord();
  PubWord(const PubWord&) = delete;
  PubWord& operator=(const PubWord&) = delete;
  // A bounded spin on the stamp; past two seconds (a stalled or profiled
  // device) a synchronisation of s, after which the word must carry the
  // stamp (else Error 7 naming 'what').
  unsigned long long wait(unsigned stamp, hipStream_t s, const char* what) const;
  // decision and float hint of a word
  static int code(unsigned long long w) { return (int)((w >> 32) & 3); }
  static double hint(unsigned long long w);
};

// Doubles in coherent host memory that kernels write directly (a residual
// history read by the host after the published word that covers it).
struct HostDoubles {
  double* h = nullptr;
  double* d = nullptr;
  size_t cap = 0;
  // grows (after synchronising s, as kernels may still hold the old
  // address) to at least n values
  void ensure(size_t n, hipStream_t s);
  HostDoubles() = default;
  ~HostDoubles();
  HostDoubles(const HostDoubles&) = delete;
  HostDoubles& operator=(const HostDoubles&) = delete;
};

class PcgSolver {
 public:
  an_pcg_options opt;
  DevCSR<double> A;
  AmgHierarchy amg;
  Reducer red;
  DevBuf<double> scal;
  PinnedScalars hs;
  DevBuf<double> r, z, p, q;
  DevBuf<float> r32, z32, p32, q32, e32;
  hipStream_t s = 0;
  bool valuesSet = false;
  double lastSetupSeconds = 0.0;
  // profiling (opt.profile): event pairs around SpMV / V-cycle launches
  struct EvPair {
    hipEvent_t a, b;
  };
  std::vector<EvPair> spmvEv, vcEv;
  size_t spmvUsed = 0, vcUsed = 0;

  PcgSolver(int n, const int* rowOff, const int* cols, long long nnz, const an_pcg_options& o,
            bool deviceInput, hipStream_t stream);
  ~PcgSolver();
  // values: device pointer; rebuild the AMG when asked or when none exists.
  void set_values_device(const double* vals, bool rebuild);
  // b, x device pointers; x holds x0.
  an_status solve(const double* b, double* x, an_solve_info* info);
  void apply_prec(const double* rr, double* zz) { amg.apply(rr, zz, s); }
  // residual 2-norms of the last solve: ||r0|| and one per PCG iteration
  // (mixed precision: ||r0|| and the FP64 residual after each refinement)
  std::vector<double> history;

 private:
  // z = M^-1 r on the solver's own r/z (FP64) or r32/z32 (FP32) vectors,
  // with r.z (and z.q) into d; replayed from a graph captured on first use
  // per (precision, target, q) and dropped when the hierarchy is rebuilt
  void precondition(bool fp32, const ApplyDots& d);
  // precision 2: z32 = M32^-1 r32 (y32 ready when yReady), then z = (double)
  // z32 with out[0] = r.z and, with qv, out[1] = z.q in FP64
  void precondition_p2(bool yReady, const double* qv, double* out);
  void drop_graphs();
  std::map<long long, hipGraphExec_t> applyGraphs;
  hipStream_t cs = nullptr;  // capture stream
  // projection onto previous solutions (opt.rhsProjection): B = A X holds
  // projM 2-norm orthonormal vectors, n doubles each
  DevBuf<double> projX, projB;
  int projM = 0;
  void proj_coef(const double* rr);
  std::pair<double, double> start_residual(const double* b, double* x);
  void project_add(const double* x);
  int pcg64(const double* b, double* x, double tol, int maxIter, double* rnormOut);
  // rn0: norm of b (the outer residual) for the speculation rule
  int pcg32(const float* b, float* x, double tol, int maxIter, double rn0, double* rnormOut);
  // Each iteration's outcome comes back as one published word (pcg_word:
  // stamp, device-side stop decision, float ||r|| hint) that the kernel
  // finishing r.r stores after it has written the FP64 ||r|| into hist
  // (coherent host memory): the host reads both without a copy.
  PubWord pub;
  unsigned pubStamp = 0;
  HostDoubles hist;
  // scal[slot .. slot + count) to the host through the published word (a
  // one-thread kernel, then a spin on its stamp); returns the host copy
  const double* fetch_scalars(int slot, int count);
  struct Published {
    int code;
    double rnHint;
  };
  Published wait_published(unsigned stamp);
  PcgPub pub_args(double tol, int idx) {
    PcgPub a;
    a.word = pub.d;
    a.stamp = ++pubStamp;
    a.tol = tol;
    a.hist = hist.d;
    a.idx = idx;
    return a;
  }
  EvPair* next_ev(std::vector<EvPair>& pool, size_t& used);
  void collect_profile(an_solve_info* inf);
};

// One-shot Jacobi-preconditioned BiCGSTAB on device data (x holds x0).
// With pinned memory (the default) each iteration's stop test runs in the
// kernel that finishes r.r and comes back through a published word; without
// (pinnedScalars = false: callers that may not allocate) the scalars are
// copied to a plain host array and the host decides, with the same result.
struct BicgWork {
  Reducer red;
  DevBuf<double> scal, dinv, r, r0, p, v, ph, s_, sh, t;
  PinnedScalars hs;
  std::unique_ptr<PubWord> pub;
  unsigned stamp = 0;
  int n = 0;
  explicit BicgWork(bool pinnedScalars = true) : hs(pinnedScalars) {
    if (pinnedScalars) pub.reset(new PubWord());
  }
  void ensure(int n, int nbSpmv, hipStream_t s);
};
an_status bicgstab_device(const DevCSR<double>& A, const double* b, double* x, int maxIter,
                          double relTol, double absTol, BicgWork& w, hipStream_t s,
                          an_solve_info* info);

// Three right-hand sides sharing one matrix (interleaved work vectors); per
// column bit-identical to bicgstab_device.  activeMask selects the columns
// to solve; evStart/evStop (optional) time the first fused SpMV.
//
// The stop tests run on the device: a one-thread step after each iteration
// updates the active-column mask that every kernel of the next iteration
// reads, and publishes whether any column is left (PubWord), so the host
// queues iteration k+1 before it knows the outcome of iteration k and never
// copies scalars back in the loop.
struct Bicg3State {
  int mask;      // columns still iterating
  int zero;      // columns with b = 0 (x set to 0)
  int its[3];    // iterations done per column
  int conv[3];   // converged per column
  double tol[3], rn[3], bnorm[3];
};
struct Bicg3Work {
  Reducer red;
  DevBuf<double> scal, dinv, x, b3, r, r0, p, v, ph, s_, sh, t;
  DevBuf<double> state;          // a Bicg3State
  Bicg3State* hs = nullptr;      // pinned copy of state, read once per solve
  PubWord pub;
  unsigned stamp = 0;
  int n = 0;
  void ensure(int n, int nbSpmv, hipStream_t s);
  Bicg3Work() = default;
  ~Bicg3Work();
  Bicg3Work(const Bicg3Work&) = delete;
  Bicg3Work& operator=(const Bicg3Work&) = delete;
};
an_status bicgstab3_device(const DevCSR<double>& A, const double* const b[3], double* const x[3],
                           int maxIter, double relTol, double absTol, Bicg3Work& w,
                           hipStream_t s, an_solve_info info[3], int activeMask = 7,
                           hipEvent_t evStart = nullptr, hipEvent_t evStop = nullptr,
                           const double* const* x0 = nullptr, const double* dinv = nullptr);

void default_pcg_options(an_pcg_options* o);

// discrete adjoint of the steady scalar transport solve (adjoint.hip)
struct DevMesh;
struct MeshView;
struct BCView;
void launch_gamma_from_cells(const MeshView& m, const double* k, double* gam, hipStream_t s);
void launch_dk_from_dgamma(const MeshView& m, const double* dgam, double* dk, hipStream_t s);
void launch_lsq_adjoint(const MeshView& m, BCView bc, const double* w, double* z, double* phibar,
                        double* phibarB, hipStream_t s);
an_status transport_adjoint(const DevMesh& d, const MeshView& mv, const double* flux,
                            const double* gam, const double* Su, const double* Sp, BCView bc,
                            int upwind, int nonOrth, int nOuter, int maxIter, double relTol,
                            double absTol, const double* phi, const double* dJdphi,
                            double* dgam, double* dbc, double* psiOut, hipStream_t s,
                            an_solve_info* info);

}  // namespace an
