// solver.hip -- PCG + aggregation AMG and Jacobi BiCGSTAB on the GPU.
//
// PCG (replaces HYPRE_ParCSRPCG + BoomerAMG, LP/src/hypre_backend.cu:54-99,
// GSV1 main.cu:2224-2273, 2617-2676): two-norm stop
// ||r|| <= max(relTol ||b||, absTol); b == 0 gives x = 0; x0 from the caller.
// All scalars (alpha, beta, dots) stay in device memory; the host reads
// ||r||^2 once per iteration through pinned memory, and the wait is
// overlapped with the next preconditioner application (which does not touch
// x), so the host round trip is off the critical path.
//
// Mixed precision: FP64 iterative refinement whose inner solve is a PCG in
// FP32 (FP32 matrix values, vectors, SpMV, smoother and AMG levels; FP64 dot
// accumulation and FP64 dense coarsest solve); the outer residual b - A x is
// FP64.
#include "solver.hpp"

#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>

namespace an {

void default_pcg_options(an_pcg_options* o) {
  o->maxIter = 4000;
  o->relTol = 0.0;
  o->absTol = 1e-12;
  o->precision = 0;
  o->coarseSize = 256;
  o->maxLevels = 25;
  o->strength = 0.25;
  o->pairPasses = 3;
  o->jacobiOmega = 0.85;
  o->innerRelTol = 1e-4;
  o->profile = 0;
  o->sweeps = 1;
  o->kcycle = 2;
  o->graphs = 0;  // replay measured no faster than launching here
  o->rhsProjection = 0;
  o->smoother = 0;
  o->chebLevels = 1;
  o->chebFraction = 0.3;
}

static constexpr int kB = 256;

PcgSolver::PcgSolver(int n, const int* rowOff, const int* cols, long long nnz,
                     const an_pcg_options& o, bool dev, hipStream_t stream)
    : opt(o), s(stream) {
  A.n = n;
  A.nnz = nnz;
  A.rowOff.alloc((size_t)n + 1);
  A.cols.alloc((size_t)nnz);
  A.vals.alloc((size_t)nnz);
  hipMemcpyKind k = dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
  AN_HIP(hipMemcpyAsync(A.rowOff.p, rowOff, sizeof(int) * ((size_t)n + 1), k, s));
  AN_HIP(hipMemcpyAsync(A.cols.p, cols, sizeof(int) * (size_t)nnz, k, s));
  A.finish_pattern(s);
  red.init(std::max(grid_for(A.rb.count, kB / 64), grid_for(n, kB)) + 1024, s);
  scal.alloc(S_COUNT);
  scal.zero(s);
  r.alloc(n);
  z.alloc(n);
  p.alloc(n);
  q.alloc(n);
  amg.opt.coarseSize = o.coarseSize;
  amg.opt.maxLevels = o.maxLevels;
  amg.opt.strength = o.strength;
  amg.opt.pairPasses = o.pairPasses;
  amg.opt.omega = o.jacobiOmega;
  amg.opt.sweeps = o.sweeps > 0 ? o.sweeps : 1;
  amg.opt.kcycle = o.kcycle > 0 ? o.kcycle : 0;
  amg.opt.smoother = o.smoother;
  amg.opt.chebLevels = o.chebLevels;
  amg.opt.chebFraction = o.chebFraction;
  AN_HIP(hipStreamSynchronize(s));
}

PcgSolver::~PcgSolver() {
  drop_graphs();
  if (cs) (void)hipStreamDestroy(cs);
  for (auto& e : spmvEv) {
    (void)hipEventDestroy(e.a);
    (void)hipEventDestroy(e.b);
  }
  for (auto& e : vcEv) {
    (void)hipEventDestroy(e.a);
    (void)hipEventDestroy(e.b);
  }
}

PcgSolver::EvPair* PcgSolver::next_ev(std::vector<EvPair>& pool, size_t& used) {
  if (!opt.profile) return nullptr;
  if (used == pool.size()) {
    EvPair e;
    AN_HIP(hipEventCreate(&e.a));
    AN_HIP(hipEventCreate(&e.b));
    pool.push_back(e);
  }
  return &pool[used++];
}

void PcgSolver::collect_profile(an_solve_info* inf) {
  inf->spmvIndexBytes = A.has16 ? 2 : 4;
  if (!opt.profile) return;
  AN_HIP(hipStreamSynchronize(s));
  double t = 0.0;
  for (size_t i = 0; i < spmvUsed; ++i) {
    float ms = 0.f;
    AN_HIP(hipEventElapsedTime(&ms, spmvEv[i].a, spmvEv[i].b));
    t += ms * 1e-3;
  }
  inf->spmvCalls = (int)spmvUsed;
  inf->spmvAvgSeconds = spmvUsed ? t / spmvUsed : 0.0;
  t = 0.0;
  for (size_t i = 0; i < vcUsed; ++i) {
    float ms = 0.f;
    AN_HIP(hipEventElapsedTime(&ms, vcEv[i].a, vcEv[i].b));
    t += ms * 1e-3;
  }
  inf->vcycleCalls = (int)vcUsed;
  inf->vcycleAvgSeconds = vcUsed ? t / vcUsed : 0.0;
}

void PcgSolver::set_values_device(const double* vals, bool rebuild) {
  AN_HIP(hipMemcpyAsync(A.vals.p, vals, sizeof(double) * (size_t)A.nnz, hipMemcpyDeviceToDevice, s));
  valuesSet = true;
  projM = 0;  // the stored solutions belong to the old matrix
  if (rebuild || !amg.built) {
    auto t0 = std::chrono::steady_clock::now();
    drop_graphs();  // the captured cycles point into the old hierarchy
    amg.setup(A, s);
    if (opt.precision == 1 || opt.precision == 2) {
      amg.make_fp32(s);
      r32.alloc(A.n);
      z32.alloc(A.n);
      p32.alloc(A.n);
      q32.alloc(A.n);
      e32.alloc(A.n);
    }
    AN_HIP(hipStreamSynchronize(s));
    lastSetupSeconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  } else {
    amg.refresh_fine(s);
  }
}

void PcgSolver::drop_graphs() {
  for (auto& g : applyGraphs) (void)hipGraphExecDestroy(g.second);
  applyGraphs.clear();
}

void PcgSolver::precondition(bool fp32, const ApplyDots& d) {
  auto launch = [&](hipStream_t st) {
    if (fp32)
      amg.apply32(r32.p, z32.p, st, &d);
    else
      amg.apply(r.p, z.p, st, &d);
  };
  if (!opt.graphs) {
    launch(s);
    return;
  }
  const long long slot = d.ra.out ? (long long)(d.ra.out - scal.p) : (long long)S_COUNT;
  const long long key = (slot << 3) | (d.yReady ? 4 : 0) | (d.q ? 2 : 0) | (fp32 ? 1 : 0);
  auto it = applyGraphs.find(key);
  if (it == applyGraphs.end()) {
    // captured on a private non-blocking stream (the solver's own stream may
    // be the null stream, which cannot capture), replayed on the solver's
    if (!cs) AN_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    hipGraph_t graph = nullptr;
    AN_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
    try {
      launch(cs);
    } catch (...) {
      (void)hipStreamEndCapture(cs, &graph);
      if (graph) (void)hipGraphDestroy(graph);
      throw;
    }
    AN_HIP(hipStreamEndCapture(cs, &graph));
    hipGraphExec_t exec = nullptr;
    const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    AN_HIP(e);
    it = applyGraphs.emplace(key, exec).first;
  }
  AN_HIP(hipGraphLaunch(it->second, s));
}

void PcgSolver::precondition_p2(bool yReady, const double* qv, double* out) {
  const ApplyDots none{nullptr, RedArgs{nullptr, nullptr}, yReady};
  precondition(true, none);  // z32 = M32^-1 r32
  z_from_fp32(A.n, z32.p, r.p, qv, z.p, red.args(out), s);
}

// The next preconditioner application is queued before the host knows
// ||r|| (hiding the round trip) unless the last two residual norms predict
// convergence in this iteration: then an application would be wasted work
// on the stream, which short solves (warm-started corrections) feel most.
static inline bool speculate(double rn, double rnPrev, double tol) {
  return !(rnPrev > 0.0 && rn * (rn / rnPrev) <= tol);
}

PubWord::PubWord() {
  AN_HIP(hipHostMalloc((void**)&h, sizeof(double) * (1 + kValues), hipHostMallocCoherent));
  __atomic_store_n(h, 0ull, __ATOMIC_RELEASE);
  AN_HIP(hipHostGetDevicePointer((void**)&d, h, 0));
  vh = reinterpret_cast<double*>(h + 1);
  vd = reinterpret_cast<double*>(d + 1);
}
PubWord::~PubWord() {
  if (h) (void)hipHostFree(h);
}

// Stamps only grow (modulo 2^30); a word at or after the awaited stamp
// answers the wait.  A later stamp is the outcome of an iteration queued
// before the host looked (the three-column BiCGSTAB runs one ahead), which
// carries the earlier one: CONTINUE only if every earlier iteration
// continued.
static inline bool stamp_reached(unsigned have, unsigned want) {
  return ((have - want) & 0x3fffffffu) < 0x20000000u;
}

unsigned long long PubWord::wait(unsigned stamp, hipStream_t s, const char* what) const {
  const unsigned want = stamp & 0x3fffffffu;
  unsigned long long w = __atomic_load_n(h, __ATOMIC_ACQUIRE);
  const auto t0 = std::chrono::steady_clock::now();
  for (unsigned i = 1; !stamp_reached((unsigned)(w >> 34), want); ++i) {
    if ((i & 4095) == 0 && std::chrono::steady_clock::now() - t0 > std::chrono::seconds(2)) {
      AN_HIP(hipStreamSynchronize(s));
      w = __atomic_load_n(h, __ATOMIC_ACQUIRE);
      if (!stamp_reached((unsigned)(w >> 34), want)) throw Error(7, what);
      break;
    }
    w = __atomic_load_n(h, __ATOMIC_ACQUIRE);
  }
  return w;
}

double PubWord::hint(unsigned long long w) {
  float v;
  const unsigned bits = (unsigned)(w & 0xffffffffu);
  std::memcpy(&v, &bits, sizeof v);
  return (double)v;
}

void HostDoubles::ensure(size_t n, hipStream_t s) {
  if (n <= cap) return;
  if (h) {
    AN_HIP(hipStreamSynchronize(s));
    (void)hipHostFree(h);
    h = d = nullptr;
    cap = 0;
  }
  const size_t c = std::max(n, (size_t)256);
  AN_HIP(hipHostMalloc((void**)&h, sizeof(double) * c, hipHostMallocCoherent));
  AN_HIP(hipHostGetDevicePointer((void**)&d, h, 0));
  cap = c;
}
HostDoubles::~HostDoubles() {
  if (h) (void)hipHostFree(h);
}

// dst[0 .. count) = src[0 .. count), then the stamp (system-scope release:
// the values are visible before the word)
__global__ void k_pub_values(const double* __restrict__ src, int count, double* dst,
                             unsigned long long* word, unsigned stamp) {
  if (threadIdx.x != 0) return;
  for (int i = 0; i < count; ++i) dst[i] = src[i];
  __hip_atomic_store(word, pcg_word(stamp, PCG_CONTINUE, 0.0f), __ATOMIC_RELEASE,
                     __HIP_MEMORY_SCOPE_SYSTEM);
}

const double* PcgSolver::fetch_scalars(int slot, int count) {
  if (count > PubWord::kValues) throw Error(7, "too many published values");
  const unsigned stamp = ++pubStamp;
  k_pub_values<<<1, 64, 0, s>>>(scal.p + slot, count, pub.vd, pub.d, stamp);
  AN_CHECK_LAUNCH();
  (void)pub.wait(stamp, s, "solver scalars were not published");
  return pub.vh;
}

PcgSolver::Published PcgSolver::wait_published(unsigned stamp) {
  const unsigned long long w = pub.wait(stamp, s, "PCG iteration was not published");
  return Published{PubWord::code(w), PubWord::hint(w)};
}

// FP64 PCG; r must already hold b - A x and scal[S_RR] = r.r.  The residual
// norm of iteration k lands in hist[k] (hist[0] = ||r0||, written here).
int PcgSolver::pcg64(const double* b, double* x, double tol, int maxIter, double* rnormOut) {
  (void)b;
  const int n = A.n;
  double rn = *rnormOut;
  if (rn <= tol) return 0;
  hist.ensure((size_t)maxIter + 2, s);
  // flexible CG (K-cycle preconditioner): alpha = r.z / p.q with r.z in
  // S_FRZ, beta = -(z.q)/(p.q) (Polak-Ribiere: z.(r_new - r_old) / r_old.z_old
  // with r_new - r_old = -alpha q); otherwise beta = rho_new / rho_old
  // precision 2: the same FP64 iteration around the FP32 cycle, always
  // flexible (an FP32 cycle is not exactly the symmetric operator CG's
  // rho_new / rho_old assumes)
  const bool p2 = opt.precision == 2;
  const bool flex = amg.flexible() || p2;
  if (p2) {
    copy_d2f(n, r.p, r32.p, s);
    precondition_p2(false, nullptr, scal.p + S_FRZ);
  } else {
    const ApplyDots d0{nullptr, red.args(scal.p + (flex ? S_FRZ : S_RHO0))};
    precondition(false, d0);  // r.z from the cycle's last kernel
  }
  pcg_direction<double>(n, p.p, z.p, scal.p, 0, 0, true, s);
  int it = 1, ret = 0;
  double rnPrev = 0.0;
  for (; it <= maxIter; ++it) {
    const int rhoOld = flex ? S_FRZ : (it & 1) ? S_RHO0 : S_RHO1;
    const int rhoNew = (it & 1) ? S_RHO1 : S_RHO0;
    EvPair* e1 = next_ev(spmvEv, spmvUsed);
    spmv<double, double>(A, p.p, q.p, p.p, scal.p, S_PQ, red, s, e1 ? e1->a : nullptr,
                         e1 ? e1->b : nullptr);
    // the update also writes the preconditioner's y = omdinv o r and
    // publishes the iteration's outcome to the host
    const PcgPub pa = pub_args(tol, it);
    if (p2)
      pcg_update_p2(n, x, r.p, p.p, q.p, scal.p, rhoOld, S_PQ, red, s, amg.lv[0].omdinv32.p,
                    r32.p, amg.lv[0].y32.p, &pa);
    else
      pcg_update<double>(n, x, r.p, p.p, q.p, scal.p, rhoOld, S_PQ, red, s, amg.lv[0].omdinv.p,
                         amg.lv[0].y.p, &pa);
    auto next_direction = [&]() {
      EvPair* e2 = next_ev(vcEv, vcUsed);
      if (e2) AN_HIP(hipEventRecord(e2->a, s));
      if (p2) {
        precondition_p2(true, q.p, scal.p + S_FRZ);
      } else {
        const ApplyDots dd = flex ? ApplyDots{q.p, red.args(scal.p + S_FRZ), true}
                                  : ApplyDots{nullptr, red.args(scal.p + rhoNew), true};
        precondition(false, dd);
      }
      if (e2) AN_HIP(hipEventRecord(e2->b, s));
      if (flex)
        pcg_direction<double>(n, p.p, z.p, scal.p, S_FZQ, S_PQ, false, s, -1.0);
      else
        pcg_direction<double>(n, p.p, z.p, scal.p, rhoNew, rhoOld, false, s);
    };
    const bool spec = speculate(rn, rnPrev, tol);
    if (spec) next_direction();  // queued while the host waits for the outcome
    const Published pb = wait_published(pa.stamp);
    rnPrev = rn;
    rn = pb.rnHint;
    if (pb.code == PCG_BREAKDOWN) {
      ret = -it;
      break;
    }
    if (pb.code == PCG_CONVERGED) {
      ret = it;
      break;
    }
    if (!spec) next_direction();
  }
  if (it > maxIter) {
    it = maxIter;
    ret = -maxIter;
  }
  // the FP64 residual norms of iterations 1..it: the host has seen every
  // iteration's word, which the kernel stored after hist[it]
  history.insert(history.end(), hist.h + 1, hist.h + 1 + it);
  if (it > 0) *rnormOut = history.back();
  return ret;
}

// FP32 inner PCG from x = 0 on A32 with FP32 AMG.
int PcgSolver::pcg32(const float* b, float* x, double tol, int maxIter, double rn0,
                     double* rnormOut) {
  const int n = A.n;
  const DevCSR<float>& A32 = amg.lv[0].A32;
  hist.ensure((size_t)maxIter + 2, s);
  fill_zero(x, sizeof(float) * (size_t)n, s);
  AN_HIP(hipMemcpyAsync(r32.p, b, sizeof(float) * (size_t)n, hipMemcpyDeviceToDevice, s));
  const bool flex = amg.flexible();
  const ApplyDots d0{nullptr, red.args(scal.p + (flex ? S_FRZ : S_RHO0))};
  precondition(true, d0);
  pcg_direction<float>(n, p32.p, z32.p, scal.p, 0, 0, true, s);
  int it = 1;
  double rn = rn0, rnPrev = 0.0;
  for (; it <= maxIter; ++it) {
    const int rhoOld = flex ? S_FRZ : (it & 1) ? S_RHO0 : S_RHO1;
    const int rhoNew = (it & 1) ? S_RHO1 : S_RHO0;
    EvPair* e1 = next_ev(spmvEv, spmvUsed);
    spmv<float, float>(A32, p32.p, q32.p, p32.p, scal.p, S_PQ, red, s, e1 ? e1->a : nullptr,
                       e1 ? e1->b : nullptr);
    const PcgPub pa = pub_args(tol, it);
    pcg_update<float>(n, x, r32.p, p32.p, q32.p, scal.p, rhoOld, S_PQ, red, s,
                      amg.lv[0].omdinv32.p, amg.lv[0].y32.p, &pa);
    auto next_direction = [&]() {
      EvPair* e2 = next_ev(vcEv, vcUsed);
      if (e2) AN_HIP(hipEventRecord(e2->a, s));
      const ApplyDots dd = flex ? ApplyDots{q32.p, red.args(scal.p + S_FRZ), true}
                                : ApplyDots{nullptr, red.args(scal.p + rhoNew), true};
      precondition(true, dd);
      if (e2) AN_HIP(hipEventRecord(e2->b, s));
      if (flex)
        pcg_direction<float>(n, p32.p, z32.p, scal.p, S_FZQ, S_PQ, false, s, -1.0);
      else
        pcg_direction<float>(n, p32.p, z32.p, scal.p, rhoNew, rhoOld, false, s);
    };
    const bool spec = speculate(rn, rnPrev, tol);
    if (spec) next_direction();
    const Published pb = wait_published(pa.stamp);
    rnPrev = rn;
    rn = pb.rnHint;
    if (pb.code != PCG_CONTINUE) break;
    if (!spec) next_direction();
  }
  if (it > maxIter) it = maxIter;
  if (it > 0) *rnormOut = hist.h[it];  // published with iteration it's word
  return it;
}

// ------------------------------------------- projection onto previous solutions
// Successive right-hand sides with one matrix (after Fischer 1998,
// "Projection techniques for iterative solution of Ax = b with successive
// right-hand sides", minimal-residual form): B = A X with the columns of B
// orthonormal in the 2-norm.  The initial guess x0 (the caller's) is moved
// to x0 + X c with c = B^T r0, which minimises ||b - A x|| over
// x0 + span X; the residual follows as r0 - B c without an SpMV.  So the
// start is never worse than x0 in the norm the stop test uses.
static constexpr int kProjMax = 16;

// out[i] = V_i . y, i < M (block partials, M per block).  M is a template
// argument so that a thread issues its M loads of a row together (a
// run-time count put each load behind its own branch and wait).
template <int M>
__global__ __launch_bounds__(kB) void k_mdot(int n, const double* __restrict__ V,
                                             const double* __restrict__ y,
                                             double* __restrict__ partial) {
  __shared__ double sh[M * (kB / 64)];
  double d[M];
#pragma unroll
  for (int i = 0; i < M; ++i) d[i] = 0.0;
  for (int j = blockIdx.x * kB + threadIdx.x; j < n; j += gridDim.x * kB) {
    const double yj = y[j];
    double v[M];
#pragma unroll
    for (int i = 0; i < M; ++i) v[i] = V[(size_t)i * n + j];
#pragma unroll
    for (int i = 0; i < M; ++i) d[i] += v[i] * yj;
  }
  block_sums<M>(d, sh);
  if (threadIdx.x == 0)
#pragma unroll
    for (int i = 0; i < M; ++i) partial[(size_t)blockIdx.x * M + i] = d[i];
}

template <int M>
static void launch_mdot(int m, int nb, hipStream_t s, int n, const double* V, const double* y,
                        double* partial) {
  if constexpr (M <= kProjMax) {
    if (m == M) {
      k_mdot<M><<<nb, kB, 0, s>>>(n, V, y, partial);
      return;
    }
    launch_mdot<M + 1>(m, nb, s, n, V, y, partial);
  } else {
    throw Error(7, "projection count out of range");
  }
}

// x += sx * X c and y += sy * B c in one pass (sums in index order); the
// same c keeps B = A X between the pair
__global__ __launch_bounds__(kB) void k_mcomb2(int n, int m, const double* __restrict__ X,
                                               const double* __restrict__ B,
                                               const double* __restrict__ c, double sx,
                                               double* __restrict__ x, double sy,
                                               double* __restrict__ y) {
  const int j = blockIdx.x * kB + threadIdx.x;
  if (j >= n) return;
  double tx = 0.0, ty = 0.0;
  for (int i = 0; i < m; ++i) {
    const double ci = c[i];
    tx += ci * X[(size_t)i * n + j];
    ty += ci * B[(size_t)i * n + j];
  }
  x[j] = x[j] + sx * tx;
  y[j] = y[j] + sy * ty;
}

// xout = xin + sx * X c and y += sy * B c (k_mcomb2's sums), with the block
// partials of y.y of the result (grid-stride, at most 1024 blocks)
__global__ __launch_bounds__(kB) void k_mcomb2_dot(int n, int m, const double* __restrict__ X,
                                                   const double* __restrict__ B,
                                                   const double* __restrict__ c, double sx,
                                                   const double* xin, double* xout, double sy,
                                                   double* __restrict__ y,
                                                   double* __restrict__ partial) {
  __shared__ double sh[kB / 64];
  double v = 0.0;
  for (int j = blockIdx.x * kB + threadIdx.x; j < n; j += gridDim.x * kB) {
    double tx = 0.0, ty = 0.0;
    for (int i = 0; i < m; ++i) {
      const double ci = c[i];
      tx += ci * X[(size_t)i * n + j];
      ty += ci * B[(size_t)i * n + j];
    }
    xout[j] = xin[j] + sx * tx;
    const double yj = y[j] + sy * ty;
    y[j] = yj;
    v += yj * yj;
  }
  double d[1] = {v};
  block_sums<1>(d, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = d[0];
}

// A solve's first scalars, finished and published in one kernel: NC 2: the
// residual SpMV's partials {r.r, b.b}; NC 1: r.r after the projection, with
// b.b from scal[bbSlot].  scal[S_RR] = r.r, scal[S_BB] = b.b, and both go
// to the host through the published word's values.
template <int NC>
__global__ __launch_bounds__(1024) void k_finalize_start(int nparts,
                                                         const double* __restrict__ partial,
                                                         double* __restrict__ scal, int bbSlot,
                                                         double* vd, unsigned long long* word,
                                                         unsigned stamp) {
  __shared__ double sh[NC * 16];
  double v[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) v[c] = 0.0;
#pragma unroll 4
  for (int i = threadIdx.x; i < nparts; i += blockDim.x)
#pragma unroll
    for (int c = 0; c < NC; ++c) v[c] += partial[(size_t)i * NC + c];
  block_sums<NC>(v, sh);
  if (threadIdx.x == 0) {
    const double rr = v[0], bb = NC == 2 ? v[NC - 1] : scal[bbSlot];
    scal[S_RR] = rr;
    scal[S_BB] = bb;
    vd[0] = rr;
    vd[1] = bb;
    __hip_atomic_store(word, pcg_word(stamp, PCG_CONTINUE, 0.0f), __ATOMIC_RELEASE,
                       __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

// X_m = w / ||Aw||, B_m = Aw / ||Aw|| with nuA = ||Aw||^2 after the
// orthogonalisation and nuB = ||A x||^2 before.  A remainder below 1e-8 of
// ||A x|| (x already in the span up to rounding) is stored as zero:
// normalising rounding noise would break B = A X.
__global__ __launch_bounds__(kB) void k_pscale(int n, const double* __restrict__ w,
                                               const double* __restrict__ Aw,
                                               const double* __restrict__ nuA,
                                               const double* __restrict__ nuB,
                                               double* __restrict__ X, double* __restrict__ B) {
  const int j = blockIdx.x * kB + threadIdx.x;
  if (j >= n) return;
  const double s = nuA[0] > 1e-16 * nuB[0] && nuA[0] > 0.0 ? 1.0 / sqrt(nuA[0]) : 0.0;
  X[j] = w[j] * s;
  B[j] = Aw[j] * s;
}

// c = B^T r into scal[S_PROJ ..]
void PcgSolver::proj_coef(const double* rr) {
  const int n = A.n;
  // at most 1024 blocks (more made the 16-component block sums and the
  // finalize cost more than the loads they hide), and no more than the
  // partial buffer (4 cap doubles) holds for projM components
  const int nb =
      std::min({grid_for(n, kB), 1024, 4 * red.cap / std::max(projM, 1)});
  launch_mdot<1>(projM, nb, s, n, projB.p, rr, red.partial.p);
  AN_CHECK_LAUNCH();
  finalize(red, nb, projM, scal.p + S_PROJ, s);
}

// The solve's start: r = b - A x with r.r and b.b in the same pass; with
// stored directions x += X c, r -= B c (c = B^T r) with r.r of the result
// in that pass; one kernel finishes scal[S_RR], scal[S_BB] and publishes
// them.  Returns {r.r, b.b} as the host reads them.
std::pair<double, double> PcgSolver::start_residual(const double* b, double* x) {
  const int n = A.n;
  const int nbR = residual_bb<double, double>(A, b, x, r.p, red, s);
  const unsigned stamp = ++pubStamp;
  if (opt.rhsProjection > 0 && projM > 0) {
    finalize(red, nbR, 2, scal.p + S_RB, s);  // {r.r, b.b} of the unprojected start
    proj_coef(r.p);
    const int nb = std::min(grid_for(n, kB), 1024);
    k_mcomb2_dot<<<nb, kB, 0, s>>>(n, projM, projX.p, projB.p, scal.p + S_PROJ, 1.0, x, x, -1.0,
                                   r.p, red.partial.p);
    AN_CHECK_LAUNCH();
    k_finalize_start<1><<<1, 1024, 0, s>>>(nb, red.partial.p, scal.p, S_RB + 1, pub.vd, pub.d,
                                           stamp);
  } else {
    k_finalize_start<2><<<1, 1024, 0, s>>>(nbR, red.partial.p, scal.p, 0, pub.vd, pub.d, stamp);
  }
  AN_CHECK_LAUNCH();
  (void)pub.wait(stamp, s, "solver start scalars were not published");
  return {pub.vh[0], pub.vh[1]};
}

// adds the solution x: Aw = A x (with ||A x||^2), w = x - X c and
// Aw -= B c with c = B^T A x (one classical Gram-Schmidt pass; w follows
// with X, so B = A X holds whatever the remaining loss of orthogonality,
// with ||Aw||^2 in the same pass), normalised by ||Aw||
void PcgSolver::project_add(const double* x) {
  const int maxV = std::min(opt.rhsProjection, kProjMax);
  if (maxV <= 0 || projM >= maxV) return;
  const int n = A.n;
  const int G = grid_for(n, kB);
  if (!projX.p) {
    projX.alloc((size_t)maxV * n);
    projB.alloc((size_t)maxV * n);
  }
  const int nbQ = spmv_self_dot<double, double>(A, x, q.p, red, s);  // q = A x
  finalize(red, nbQ, 1, scal.p + S_NU + 1, s);
  const double* w = x;
  const double* nuA = scal.p + S_NU + 1;
  if (projM > 0) {
    proj_coef(q.p);
    const int nb = std::min(G, 1024);
    k_mcomb2_dot<<<nb, kB, 0, s>>>(n, projM, projX.p, projB.p, scal.p + S_PROJ, -1.0, x, p.p,
                                   -1.0, q.p, red.partial.p);
    AN_CHECK_LAUNCH();
    finalize(red, nb, 1, scal.p + S_NU, s);
    w = p.p;
    nuA = scal.p + S_NU;
  }
  k_pscale<<<G, kB, 0, s>>>(n, w, q.p, nuA, scal.p + S_NU + 1, projX.p + (size_t)projM * n,
                            projB.p + (size_t)projM * n);
  AN_CHECK_LAUNCH();
  ++projM;
}

an_status PcgSolver::solve(const double* b, double* x, an_solve_info* info) {
  if (!valuesSet) throw Error(1, "solver has no matrix values");
  const int n = A.n;
  spmvUsed = vcUsed = 0;
  auto t0 = std::chrono::steady_clock::now();
  // ||b|| and the initial residual in one round trip (the residual is
  // unused when b == 0)
  const std::pair<double, double> rb = start_residual(b, x);
  const double bnorm = std::sqrt(rb.second);
  const double rn0 = std::sqrt(rb.first);
  history.clear();
  history.push_back(rn0);
  an_solve_info inf{};
  inf.levels = amg.levels();
  inf.setupSeconds = lastSetupSeconds;
  an_status st = AN_OK;
  if (bnorm == 0.0) {
    fill_zero(x, sizeof(double) * (size_t)n, s);
    AN_HIP(hipStreamSynchronize(s));
    inf.converged = 1;
  } else {
    const double tol = std::fmax(opt.relTol * bnorm, opt.absTol);
    double rn = rn0;
    if (opt.precision != 1) {
      int it = pcg64(b, x, tol, opt.maxIter, &rn);
      inf.iterations = it < 0 ? -it : it;
      inf.converged = (it >= 0 && rn <= tol) ? 1 : 0;
    } else {
      int total = 0, steps = 0;
      while (rn > tol && total < opt.maxIter) {
        copy_d2f(n, r.p, r32.p, s);
        // inner solve to innerRelTol relative reduction, not below the
        // outer absolute target
        double innerTol = std::fmax(opt.innerRelTol * rn, 0.5 * tol);
        float* rin = r32.p;
        double rin_n = 0.0;
        // the FP32 PCG overwrites r32; keep the right-hand side in q32
        AN_HIP(hipMemcpyAsync(q32.p, rin, sizeof(float) * (size_t)n, hipMemcpyDeviceToDevice, s));
        int it = pcg32(q32.p, e32.p, innerTol, opt.maxIter - total, rn, &rin_n);
        total += it;
        axpy_f2d(n, x, e32.p, s);
        residual<double, double>(A, b, x, r.p, scal.p, red, s);
        double rnNew = std::sqrt(*fetch_scalars(S_RR, 1));
        history.push_back(rnNew);
        ++steps;
        if (!std::isfinite(rnNew) || (steps > 3 && rnNew > 0.9 * rn)) {
          rn = rnNew;
          break;
        }
        rn = rnNew;
      }
      inf.iterations = total;
      inf.refinementSteps = steps;
      inf.converged = rn <= tol ? 1 : 0;
    }
    inf.finalResNorm = rn;
    inf.finalRelResNorm = rn / bnorm;
    if (!inf.converged) st = AN_ENOCONV;
  }
  if (bnorm != 0.0) project_add(x);
  inf.solveSeconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  collect_profile(&inf);
  if (info) *info = inf;
  return st;
}

// ---------------------------------------------------------- BiCGSTAB ---
void BicgWork::ensure(int nn, int nbSpmv, hipStream_t st) {
  if (n == nn && scal.p && red.cap >= nbSpmv + 16) return;
  n = nn;
  red.init(nbSpmv > 2048 ? nbSpmv + 16 : 2048, st);
  scal.alloc(S_COUNT);
  dinv.alloc(n);
  r.alloc(n);
  r0.alloc(n);
  p.alloc(n);
  v.alloc(n);
  ph.alloc(n);
  s_.alloc(n);
  sh.alloc(n);
  t.alloc(n);
}

__device__ double bsum(double v, double* sh) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) t += sh[i];
  return t;
}

// block partials of one or two sums; finalize() reduces them in index order
__device__ __forceinline__ void store1(double a, double* __restrict__ partial) {
  if (threadIdx.x == 0) partial[blockIdx.x] = a;
}
__device__ __forceinline__ void store2(double a, double b, double* __restrict__ partial) {
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = a;
    partial[2 * blockIdx.x + 1] = b;
  }
}

__global__ void k_dinv(int n, const int* __restrict__ ro, const int* __restrict__ ci,
                       const double* __restrict__ va, double* __restrict__ dinv) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double d = 0.0;
  for (int k = ro[i]; k < ro[i + 1]; ++k)
    if (ci[k] == i) d = va[k];
  dinv[i] = (d != 0.0) ? 1.0 / d : 1.0;
}

// p = r + beta (p - omega v); ph = dinv p
__global__ __launch_bounds__(kB) void k_bicg_p(int n, const double* __restrict__ r,
                                               double* __restrict__ p, const double* __restrict__ v,
                                               const double* __restrict__ dinv,
                                               double* __restrict__ ph,
                                               const double* __restrict__ sc, int rhoNew,
                                               int rhoOld, int first) {
  const double beta = first ? 0.0 : (sc[rhoNew] / sc[rhoOld]) * (sc[S_AL] / sc[S_OM]);
  const double om = first ? 0.0 : sc[S_OM];
  for (int i = blockIdx.x * kB + threadIdx.x; i < n; i += gridDim.x * kB) {
    double pi = first ? r[i] : r[i] + beta * (p[i] - om * v[i]);
    p[i] = pi;
    ph[i] = dinv[i] * pi;
  }
}

// alpha = rhoNew / r0v; s = r - alpha v; sh = dinv s; ss = s.s
__global__ __launch_bounds__(kB) void k_bicg_s(int n, const double* __restrict__ r,
                                               const double* __restrict__ v,
                                               const double* __restrict__ dinv,
                                               double* __restrict__ s_, double* __restrict__ sh,
                                               double* __restrict__ sc, int rhoNew,
                                               double* __restrict__ partial) {
  __shared__ double shm[kB / 64];
  const double alpha = sc[rhoNew] / sc[S_R0V];
  double acc = 0.0;
  for (int i = blockIdx.x * kB + threadIdx.x; i < n; i += gridDim.x * kB) {
    double si = r[i] - alpha * v[i];
    s_[i] = si;
    sh[i] = dinv[i] * si;
    acc += si * si;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) sc[S_AL] = alpha;  // read by later kernels only
  double b = bsum(acc, shm);
  store1(b, partial);
}

// omega = ts/tt; if ||s|| <= tol: x += alpha ph, r = s; else x += alpha ph +
// omega sh, r = s - omega t.  ph = dinv o p and sh = dinv o s are formed
// here from p and s (the products k_bicg_p and k_bicg_s stored for the
// SpMV, so the same bits) rather than read back.  Block partials
// of r.r and of the next iteration's rho = r0.r, two per block; rho is
// accumulated in the element order and tree of dot() over the same grid,
// so it has the bits of a separate dot(r0, r).
__global__ __launch_bounds__(kB) void k_bicg_x(int n, double* __restrict__ x,
                                               double* __restrict__ r,
                                               const double* __restrict__ p,
                                               const double* __restrict__ s_,
                                               const double* __restrict__ t,
                                               const double* __restrict__ dinv,
                                               const double* __restrict__ r0,
                                               double* __restrict__ sc, double tol,
                                               double* __restrict__ partial) {
  __shared__ double shm[kB / 64];
  const double alpha = sc[S_AL];
  const double tt = sc[S_TT];
  const double omega = (tt > 0.0) ? sc[S_TS] / tt : 0.0;
  const bool convS = sqrt(sc[S_SS]) <= tol;
  double acc = 0.0, rho = 0.0;
  for (int i = blockIdx.x * kB + threadIdx.x; i < n; i += gridDim.x * kB) {
    const double di = dinv[i];
    double ri;
    if (convS) {
      x[i] += alpha * (di * p[i]);
      ri = s_[i];
    } else {
      x[i] += alpha * (di * p[i]) + omega * (di * s_[i]);
      ri = s_[i] - omega * t[i];
    }
    r[i] = ri;
    acc += ri * ri;
    rho += r0[i] * ri;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) sc[S_OM] = omega;
  const double b = bsum(acc, shm);
  const double c = bsum(rho, shm);
  store2(b, c, partial);
}

// finalize of r.r and of the next rho (k_finalize_nc<2>'s loads and tree,
// so the bits of two separate finalizes) into scal[S_RR] and scal[rhoNext],
// then, with a word, the iteration's stop test -- converged if ||s|| or
// ||r|| <= tol, breakdown if rho, r0.v or omega = 0 or ||r|| is not
// finite -- published with the FP64 ||r|| in the word's first value
__global__ __launch_bounds__(1024) void k_bicg_fin_decide(int nparts,
                                                          const double* __restrict__ partial,
                                                          double* __restrict__ scal, int rhoNew,
                                                          int rhoNext, double tol, double* val,
                                                          unsigned long long* word,
                                                          unsigned stamp) {
  __shared__ double sh[2 * 16];
  double v[2] = {0.0, 0.0};
#pragma unroll 4
  for (int i = threadIdx.x; i < nparts; i += blockDim.x) {
    v[0] += partial[2 * (size_t)i];
    v[1] += partial[2 * (size_t)i + 1];
  }
  block_sums<2>(v, sh);
  if (threadIdx.x != 0) return;
  scal[S_RR] = v[0];
  scal[rhoNext] = v[1];
  if (!word) return;
  const double rn = sqrt(v[0]);
  const int code = (sqrt(scal[S_SS]) <= tol || rn <= tol) ? PCG_CONVERGED
                   : (scal[rhoNew] == 0.0 || scal[S_R0V] == 0.0 || scal[S_OM] == 0.0 ||
                      !isfinite(rn))
                       ? PCG_BREAKDOWN
                       : PCG_CONTINUE;
  val[0] = rn;
  __hip_atomic_store(word, pcg_word(stamp, code, (float)rn), __ATOMIC_RELEASE,
                     __HIP_MEMORY_SCOPE_SYSTEM);
}

an_status bicgstab_device(const DevCSR<double>& A, const double* b, double* x, int maxIter,
                          double relTol, double absTol, BicgWork& w, hipStream_t s,
                          an_solve_info* info) {
  const int n = A.n;
  auto t0 = std::chrono::steady_clock::now();
  w.ensure(n, grid_for(A.rb.count, kB / 64), s);
  const int G = std::min(grid_for(n, kB), 1024);
  an_solve_info inf{};
  k_dinv<<<grid_for(n, kB), kB, 0, s>>>(n, A.rowOff.p, A.cols.p, A.vals.p, w.dinv.p);
  AN_CHECK_LAUNCH();
  dot<double>(n, b, b, w.scal.p, S_BB, w.red, s);
  residual<double, double>(A, b, x, w.r.p, w.scal.p, w.red, s);
  PubWord* pw = w.pub.get();
  double bb, rr;
  if (pw) {  // S_RR .. S_BB through the word
    const unsigned st0 = ++w.stamp;
    k_pub_values<<<1, 64, 0, s>>>(w.scal.p + S_RR, S_BB - S_RR + 1, pw->vd, pw->d, st0);
    AN_CHECK_LAUNCH();
    (void)pw->wait(st0, s, "BiCGSTAB scalars were not published");
    rr = pw->vh[0];
    bb = pw->vh[S_BB - S_RR];
  } else {
    AN_HIP(hipMemcpyAsync(w.hs.h, w.scal.p, sizeof(double) * S_COUNT, hipMemcpyDeviceToHost, s));
    AN_HIP(hipStreamSynchronize(s));
    rr = w.hs.h[S_RR];
    bb = w.hs.h[S_BB];
  }
  const double bnorm = std::sqrt(bb);
  double rn = std::sqrt(rr);
  an_status st = AN_OK;
  if (bnorm == 0.0) {
    fill_zero(x, sizeof(double) * (size_t)n, s);
    AN_HIP(hipStreamSynchronize(s));
    inf.converged = 1;
    if (info) *info = inf;
    return AN_OK;
  }
  const double tol = std::fmax(relTol * bnorm, absTol);
  int it = 0;
  bool conv = rn <= tol;
  if (!conv) {
    AN_HIP(hipMemcpyAsync(w.r0.p, w.r.p, sizeof(double) * (size_t)n, hipMemcpyDeviceToDevice, s));
    // rho of the first iteration; each later one comes out of k_bicg_x
    dot<double>(n, w.r0.p, w.r.p, w.scal.p, S_RHO1, w.red, s);
    for (it = 1; it <= maxIter; ++it) {
      const int rhoNew = (it & 1) ? S_RHO1 : S_RHO0;
      const int rhoOld = (it & 1) ? S_RHO0 : S_RHO1;
      k_bicg_p<<<G, kB, 0, s>>>(n, w.r.p, w.p.p, w.v.p, w.dinv.p, w.ph.p, w.scal.p, rhoNew, rhoOld,
                                it == 1 ? 1 : 0);
      AN_CHECK_LAUNCH();
      spmv<double, double>(A, w.ph.p, w.v.p, w.r0.p, w.scal.p, S_R0V, w.red, s);
      k_bicg_s<<<G, kB, 0, s>>>(n, w.r.p, w.v.p, w.dinv.p, w.s_.p, w.sh.p, w.scal.p, rhoNew,
                                w.red.partial.p);
      AN_CHECK_LAUNCH();
      finalize(w.red, G, 1, w.scal.p + S_SS, s);
      const int nb = spmv_dot_self(A, w.sh.p, w.t.p, w.s_.p, w.red, s);
      finalize(w.red, nb, 2, w.scal.p + S_TS, s);  // S_TS, S_TT
      k_bicg_x<<<G, kB, 0, s>>>(n, x, w.r.p, w.p.p, w.s_.p, w.t.p, w.dinv.p, w.r0.p, w.scal.p,
                                tol, w.red.partial.p);
      AN_CHECK_LAUNCH();
      if (pw) {
        const unsigned sti = ++w.stamp;
        k_bicg_fin_decide<<<1, 1024, 0, s>>>(G, w.red.partial.p, w.scal.p, rhoNew, rhoOld, tol,
                                             pw->vd, pw->d, sti);
        AN_CHECK_LAUNCH();
        const unsigned long long word = pw->wait(sti, s, "BiCGSTAB iteration was not published");
        rn = pw->vh[0];
        if (PubWord::code(word) == PCG_CONVERGED) {
          conv = true;
          break;
        }
        if (PubWord::code(word) == PCG_BREAKDOWN) break;
        continue;
      }
      k_bicg_fin_decide<<<1, 1024, 0, s>>>(G, w.red.partial.p, w.scal.p, rhoNew, rhoOld, tol,
                                           nullptr, nullptr, 0);
      AN_CHECK_LAUNCH();
      AN_HIP(hipMemcpyAsync(w.hs.h, w.scal.p, sizeof(double) * S_COUNT, hipMemcpyDeviceToHost, s));
      AN_HIP(hipStreamSynchronize(s));
      const double* h = w.hs.h;
      rn = std::sqrt(h[S_RR]);
      if (std::sqrt(h[S_SS]) <= tol || rn <= tol) {
        conv = true;
        break;
      }
      if (h[rhoNew] == 0.0 || h[S_R0V] == 0.0 || h[S_OM] == 0.0 || !std::isfinite(rn)) break;
    }
    if (it > maxIter) it = maxIter;
  }
  inf.iterations = it;
  inf.converged = conv ? 1 : 0;
  inf.finalResNorm = rn;
  inf.finalRelResNorm = rn / bnorm;
  inf.solveSeconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  inf.spmvIndexBytes = A.has16 ? 2 : 4;
  if (!conv) st = AN_ENOCONV;
  if (info) *info = inf;
  return st;
}

}  // namespace an
