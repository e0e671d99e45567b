// batch.hip -- many independent PCG + aggregation-AMG solves as one
// block-diagonal system (the batched MMS solves: many small, uneven Poisson
// problems where one solve per launch sequence would leave the GPU idle).
//
// * One CSR holds every problem (rows and columns offset per problem); the
//   AMG hierarchy of the block-diagonal matrix never couples problems, so it
//   is every problem's own hierarchy, and one V-cycle launch sequence serves
//   all of them.
// * Row blocks never cross a problem boundary; every wave reduces its block
//   to one partial and a per-problem finalize adds the partials of that
//   problem in block order, so each problem has its own deterministic
//   Krylov scalars, stop test and iteration count.  A converged problem is
//   switched off by a device-side flag: its waves skip the Krylov kernels.
// * The PCG iteration (SpMV, updates, per-problem reductions, stop test,
//   V-cycle) is captured once in a HIP graph of several iterations and
//   replayed until no problem is active; the host reads one counter per
//   replay.
// * Autotuner: the block height (64, 32 or 16 rows per wave) is chosen by
//   timing the batched SpMV on the actual matrix when values are set.
// * Precision: FP64 PCG, or FP64 iterative refinement around an FP32 PCG
//   (FP32 matrix, vectors and V-cycle; FP64 residuals and reductions).
#include "batch.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>

namespace an {

static constexpr int kB = 256;
static constexpr int kW = kB / 64;
static constexpr int kGraphIters = 4;  // iterations per captured graph (even)

__device__ __forceinline__ double bwsum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

__device__ __forceinline__ bool wave_on(int gw, int nW, const int* wPid, const int* act) {
  return gw < nW && (act == nullptr || act[wPid[gw]] != 0);
}

// MODE 0: y = A x; 1: y = b - A x.  DOT 1: dw.y, 2: y.y -> partial[wave]
template <typename T, int MODE, int DOT>
__global__ __launch_bounds__(kB) void kb_spmv(int nW, const int2* __restrict__ rbs,
                                              const int* __restrict__ wPid,
                                              const int* __restrict__ act,
                                              const int* __restrict__ rowOff,
                                              const int* __restrict__ cols,
                                              const T* __restrict__ vals, const T* __restrict__ x,
                                              const double* __restrict__ b, T* __restrict__ y,
                                              const T* __restrict__ dw,
                                              double* __restrict__ partial) {
  __shared__ double prod[kW][kStreamCap];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int gw = xcd_block(gridDim.x) * kW + w;
  const bool on = wave_on(gw, nW, wPid, act);
  WaveRows q{0, 0, 0, 0, false};
  if (on) q = wave_rows(nW, rbs, gw);
  const int row = q.r0 + lane;
  int kb = 0, ke = 0;
  if (row < q.r1) {
    kb = rowOff[row];
    ke = rowOff[row + 1];
  }
  if (q.stream && q.k1 > q.k0)
    stage_products(q, lane, cols, vals, [&](int j) { return (double)x[j]; }, prod[w]);
  wave_sync();
  double dv = 0.0;
  if (row < q.r1) {
    double s = 0.0;
    if (q.stream)
      for (int k = kb; k < ke; ++k) s += prod[w][k - q.k0];
    else
      for (int k = kb; k < ke; ++k) s += (double)vals[k] * (double)x[cols[k]];
    const T v = (MODE == 1) ? (T)(b[row] - s) : (T)s;
    y[row] = v;
    if (DOT == 1) dv = (double)dw[row] * (double)v;
    if (DOT == 2) dv = (double)v * (double)v;
  }
  if (DOT) {
    dv = bwsum(dv);
    if (lane == 0 && gw < nW) partial[gw] = dv;
  }
}

// rows of the wave's block: lane -> row (rows <= 64 per block)
__device__ __forceinline__ int wave_row(int nW, const int2* rbs, int gw, int lane, bool on) {
  if (!on) return -1;
  const int r0 = rbs[gw].x, r1 = rbs[gw + 1].x;
  return (r0 + lane < r1) ? r0 + lane : -1;
}

template <typename T>
__global__ __launch_bounds__(kB) void kb_dot(int nW, const int2* __restrict__ rbs,
                                             const int* __restrict__ wPid,
                                             const int* __restrict__ act, const T* __restrict__ a,
                                             const T* __restrict__ b,
                                             double* __restrict__ partial) {
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int gw = xcd_block(gridDim.x) * kW + w;
  const int i = wave_row(nW, rbs, gw, lane, wave_on(gw, nW, wPid, act));
  double v = (i >= 0) ? (double)a[i] * (double)b[i] : 0.0;
  v = bwsum(v);
  if (lane == 0 && gw < nW) partial[gw] = v;
}

// x += alpha p ; r -= alpha q ; partial = r.r   (alpha = num / den per problem)
template <typename T>
__global__ __launch_bounds__(kB) void kb_update(int nW, const int2* __restrict__ rbs,
                                                const int* __restrict__ wPid,
                                                const int* __restrict__ act, T* __restrict__ x,
                                                T* __restrict__ r, const T* __restrict__ p,
                                                const T* __restrict__ q,
                                                const double* __restrict__ scal, int num, int den,
                                                double* __restrict__ partial) {
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int gw = xcd_block(gridDim.x) * kW + w;
  const bool on = wave_on(gw, nW, wPid, act);
  const int i = wave_row(nW, rbs, gw, lane, on);
  double v = 0.0;
  if (i >= 0) {
    const double* sc = scal + (size_t)S_COUNT * wPid[gw];
    const double alpha = sc[num] / sc[den];
    x[i] = (T)((double)x[i] + alpha * (double)p[i]);
    const T ri = (T)((double)r[i] - alpha * (double)q[i]);
    r[i] = ri;
    v = (double)ri * (double)ri;
  }
  v = bwsum(v);
  if (lane == 0 && gw < nW) partial[gw] = v;
}

// p = z + beta p (beta = num / den), or p = z
template <typename T>
__global__ __launch_bounds__(kB) void kb_dir(int nW, const int2* __restrict__ rbs,
                                             const int* __restrict__ wPid,
                                             const int* __restrict__ act, T* __restrict__ p,
                                             const T* __restrict__ z,
                                             const double* __restrict__ scal, int num, int den,
                                             int first) {
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int gw = xcd_block(gridDim.x) * kW + w;
  const int i = wave_row(nW, rbs, gw, lane, wave_on(gw, nW, wPid, act));
  if (i < 0) return;
  if (first) {
    p[i] = z[i];
  } else {
    const double* sc = scal + (size_t)S_COUNT * wPid[gw];
    const double beta = sc[num] / sc[den];
    p[i] = (T)((double)z[i] + beta * (double)p[i]);
  }
}

// scal[p][slot] = sum of problem p's wave partials in block order
__global__ __launch_bounds__(kB) void kb_finalize(const int* __restrict__ wOff,
                                                  const double* __restrict__ partial,
                                                  double* __restrict__ scal, int slot) {
  __shared__ double sh[kW];
  const int p = blockIdx.x;
  double v = 0.0;
  for (int i = wOff[p] + threadIdx.x; i < wOff[p + 1]; i += kB) v += partial[i];
  v = bwsum(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) sh[w] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int k = 0; k < kW; ++k) t += sh[k];
    scal[(size_t)S_COUNT * p + slot] = t;
  }
}

// stop test after the update (pcg64's): converged, breakdown or out of
// iterations -> inactive.  nAct = number of problems still active.
__global__ __launch_bounds__(kB) void kb_check(int nP, const double* __restrict__ scal,
                                               const double* __restrict__ tol2,
                                               int* __restrict__ act, int* __restrict__ its,
                                               int* __restrict__ nAct, int maxIter) {
  __shared__ int cnt[kW];
  int c = 0;
  for (int p = threadIdx.x; p < nP; p += kB) {
    if (!act[p]) continue;
    const int it = its[p] + 1;
    its[p] = it;
    const double* sc = scal + (size_t)S_COUNT * p;
    const double rr = sc[S_RR], pq = sc[S_PQ];
    const bool stop = !(rr > tol2[p]) || !(pq > 0.0) || !isfinite(rr) || it >= maxIter;
    if (stop)
      act[p] = 0;
    else
      ++c;
  }
  for (int off = 32; off > 0; off >>= 1) c += __shfl_down(c, off, 64);
  if ((threadIdx.x & 63) == 0) cnt[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) nAct[0] = cnt[0] + cnt[1] + cnt[2] + cnt[3];
}

__global__ void kb_shift(int* __restrict__ a, int count, int add) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) a[i] += add;
}

// ------------------------------------------------------------------ host
BatchSolver::BatchSolver(int nProblems, const std::vector<int>& rows,
                         const std::vector<const int*>& ro, const std::vector<const int*>& ci,
                         bool dev, const an_pcg_options& o)
    : nP(nProblems), opt(o) {
  if (nP <= 0) throw Error(AN_EINVAL, "batch needs at least one problem");
  AN_HIP(hipStreamCreate(&s));
  const hipMemcpyKind k = dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
  pOff.assign(nP + 1, 0);
  pNnzOff.assign(nP + 1, 0);
  for (int p = 0; p < nP; ++p) {
    if (rows[p] <= 0 || !ro[p] || !ci[p]) throw Error(AN_EINVAL, "empty batch problem");
    int last = 0;
    AN_HIP(hipMemcpy(&last, ro[p] + rows[p], sizeof(int), dev ? hipMemcpyDeviceToHost
                                                                : hipMemcpyHostToHost));
    pOff[p + 1] = pOff[p] + rows[p];
    pNnzOff[p + 1] = pNnzOff[p] + last;
  }
  n = pOff[nP];
  nnz = pNnzOff[nP];
  A.n = n;
  A.nnz = nnz;
  A.rowOff.alloc((size_t)n + 1);
  A.cols.alloc((size_t)nnz);
  A.vals.alloc((size_t)nnz);
  for (int p = 0; p < nP; ++p) {
    AN_HIP(hipMemcpyAsync(A.rowOff.p + pOff[p], ro[p], sizeof(int) * (size_t)rows[p], k, s));
    AN_HIP(hipMemcpyAsync(A.cols.p + pNnzOff[p], ci[p],
                          sizeof(int) * (size_t)(pNnzOff[p + 1] - pNnzOff[p]), k, s));
    kb_shift<<<grid_for(rows[p], kB), kB, 0, s>>>(A.rowOff.p + pOff[p], rows[p], (int)pNnzOff[p]);
    kb_shift<<<grid_for(pNnzOff[p + 1] - pNnzOff[p], kB), kB, 0, s>>>(
        A.cols.p + pNnzOff[p], (int)(pNnzOff[p + 1] - pNnzOff[p]), pOff[p]);
  }
  const int nnzI = (int)nnz;
  AN_HIP(hipMemcpyAsync(A.rowOff.p + n, &nnzI, sizeof(int), hipMemcpyHostToDevice, s));
  AN_CHECK_LAUNCH();
  A.finish_pattern(s);
  amg.opt.coarseSize = o.coarseSize;
  amg.opt.maxLevels = o.maxLevels;
  amg.opt.strength = o.strength;
  amg.opt.pairPasses = o.pairPasses;
  amg.opt.omega = o.jacobiOmega;
  amg.opt.sweeps = o.sweeps > 0 ? o.sweeps : 1;
  amg.opt.kcycle = 0;  // the K-cycle's inner products would mix the problems
  scal.alloc((size_t)S_COUNT * nP);
  scal.zero(s);
  tol2.alloc((size_t)2 * nP);  // outer, inner
  act.alloc(nP);
  its.alloc(nP);
  nAct.alloc(1);
  for (DevBuf<double>* v : {&bB, &xB, &r, &z, &p, &q}) v->alloc((size_t)n);
  AN_HIP(hipHostMalloc((void**)&hAct, sizeof(int) * (size_t)(nP + 1)));
  AN_HIP(hipHostMalloc((void**)&hScal, sizeof(double) * (size_t)S_COUNT * nP));
  AN_HIP(hipStreamSynchronize(s));
}

BatchSolver::~BatchSolver() {
  if (g64) (void)hipGraphExecDestroy(g64);
  if (g32) (void)hipGraphExecDestroy(g32);
  if (hAct) (void)hipHostFree(hAct);
  if (hScal) (void)hipHostFree(hScal);
  if (s) (void)hipStreamDestroy(s);
}

void BatchSolver::build_layout(int rows, BatchLayout& L) {
  std::vector<int> ro((size_t)n + 1);
  AN_HIP(hipMemcpyAsync(ro.data(), A.rowOff.p, sizeof(int) * ((size_t)n + 1),
                        hipMemcpyDeviceToHost, s));
  AN_HIP(hipStreamSynchronize(s));
  const int cap = kStreamPer * rows;  // <= 12 staged nonzeros per lane
  std::vector<int2> st;
  std::vector<int> pid, woff(nP + 1, 0);
  for (int p = 0; p < nP; ++p) {
    woff[p] = (int)st.size();
    int r = pOff[p];
    while (r < pOff[p + 1]) {
      const int s0 = r++;
      while (r < pOff[p + 1] && r - s0 < rows && ro[r + 1] - ro[s0] <= cap) ++r;
      st.push_back(make_int2(s0, ro[s0]));
      pid.push_back(p);
    }
  }
  woff[nP] = (int)st.size();
  L.nW = (int)st.size();
  st.push_back(make_int2(n, ro[n]));
  L.rows = rows;
  L.rb.count = L.nW;
  L.rb.start.upload(st, s);
  L.wPid.upload(pid, s);
  L.wOff.upload(woff, s);
  AN_HIP(hipStreamSynchronize(s));
}

double BatchSolver::time_layout(const BatchLayout& L) {
  DevBuf<double> part((size_t)L.nW + 1);
  const int G = grid_for(L.nW, kW);
  auto launch = [&]() {
    kb_spmv<double, 0, 1><<<G, kB, 0, s>>>(L.nW, L.rb.start.p, L.wPid.p, nullptr, A.rowOff.p,
                                           A.cols.p, A.vals.p, p.p, nullptr, q.p, p.p, part.p);
  };
  for (int i = 0; i < 3; ++i) launch();
  hipEvent_t e0, e1;
  AN_HIP(hipEventCreate(&e0));
  AN_HIP(hipEventCreate(&e1));
  AN_HIP(hipEventRecord(e0, s));
  const int reps = 10;
  for (int i = 0; i < reps; ++i) launch();
  AN_HIP(hipEventRecord(e1, s));
  AN_HIP(hipEventSynchronize(e1));
  AN_CHECK_LAUNCH();
  float ms = 0.f;
  AN_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return ms * 1e-3 / reps;
}

void BatchSolver::set_values(const std::vector<const double*>& values, bool dev) {
  const hipMemcpyKind k = dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
  for (int pp = 0; pp < nP; ++pp)
    AN_HIP(hipMemcpyAsync(A.vals.p + pNnzOff[pp], values[pp],
                          sizeof(double) * (size_t)(pNnzOff[pp + 1] - pNnzOff[pp]), k, s));
  amg.setup(A, s);
  if (opt.precision == 1) {
    amg.make_fp32(s);
    for (DevBuf<float>* v : {&r32, &z32, &p32, &q32, &e32}) v->alloc((size_t)n);
  }
  // autotuner: block height for the batched kernels
  fill_zero(p.p, sizeof(double) * (size_t)n, s);
  tuneRows = {64, 32, 16};
  tuneSeconds.assign(tuneRows.size(), 0.0);
  double best = 1e300;
  int bestRows = 64;
  for (size_t v = 0; v < tuneRows.size(); ++v) {
    BatchLayout L;
    build_layout(tuneRows[v], L);
    tuneSeconds[v] = time_layout(L);
    if (tuneSeconds[v] < best) {
      best = tuneSeconds[v];
      bestRows = tuneRows[v];
    }
  }
  build_layout(bestRows, lay);
  partial.alloc((size_t)lay.nW + 1);
  if (g64) (void)hipGraphExecDestroy(g64);
  if (g32) (void)hipGraphExecDestroy(g32);
  g64 = g32 = nullptr;
  valuesSet = true;
}

void BatchSolver::finalize(int slot) {
  kb_finalize<<<nP, kB, 0, s>>>(lay.wOff.p, partial.p, scal.p, slot);
  AN_CHECK_LAUNCH();
}

void BatchSolver::read_scalars() {
  AN_HIP(hipMemcpyAsync(hScal, scal.p, sizeof(double) * (size_t)S_COUNT * nP,
                        hipMemcpyDeviceToHost, s));
  AN_HIP(hipStreamSynchronize(s));
}

void BatchSolver::residual64() {
  kb_spmv<double, 1, 2><<<grid_for(lay.nW, kW), kB, 0, s>>>(
      lay.nW, lay.rb.start.p, lay.wPid.p, nullptr, A.rowOff.p, A.cols.p, A.vals.p, xB.p, bB.p,
      r.p, nullptr, partial.p);
  AN_CHECK_LAUNCH();
  finalize(S_RR);
}

static void vcycle_apply(AmgHierarchy& amg, const double* r, double* z, hipStream_t s) {
  amg.apply(r, z, s);
}
static void vcycle_apply(AmgHierarchy& amg, const float* r, float* z, hipStream_t s) {
  amg.apply32(r, z, s);
}

// z = M r ; rho0 = r.z ; p = z   (r holds the right-hand side, x = 0 or x0)
template <typename T>
void BatchSolver::pcg_prologue(const DevCSR<T>&, T* xv, T* rv, T* zv, T* pv) {
  (void)xv;
  const int G = grid_for(lay.nW, kW);
  vcycle_apply(amg, rv, zv, s);
  kb_dot<T><<<G, kB, 0, s>>>(lay.nW, lay.rb.start.p, lay.wPid.p, act.p, rv, zv, partial.p);
  AN_CHECK_LAUNCH();
  finalize(S_RHO0);
  kb_dir<T><<<G, kB, 0, s>>>(lay.nW, lay.rb.start.p, lay.wPid.p, act.p, pv, zv, scal.p, 0, 0, 1);
  AN_CHECK_LAUNCH();
}

template <typename T>
void BatchSolver::pcg_iteration(const DevCSR<T>& M, int it, T* xv, T* rv, T* zv, T* pv, T* qv) {
  const int rhoOld = (it & 1) ? S_RHO0 : S_RHO1;
  const int rhoNew = (it & 1) ? S_RHO1 : S_RHO0;
  const int G = grid_for(lay.nW, kW);
  const int2* rbs = lay.rb.start.p;
  kb_spmv<T, 0, 1><<<G, kB, 0, s>>>(lay.nW, rbs, lay.wPid.p, act.p, M.rowOff.p, M.cols.p,
                                    M.vals.p, pv, nullptr, qv, pv, partial.p);
  AN_CHECK_LAUNCH();
  finalize(S_PQ);
  kb_update<T><<<G, kB, 0, s>>>(lay.nW, rbs, lay.wPid.p, act.p, xv, rv, pv, qv, scal.p, rhoOld,
                                S_PQ, partial.p);
  AN_CHECK_LAUNCH();
  finalize(S_RR);
  const double* t2 = tol2.p + (sizeof(T) == 4 ? nP : 0);
  kb_check<<<1, kB, 0, s>>>(nP, scal.p, t2, act.p, its.p, nAct.p, opt.maxIter);
  vcycle_apply(amg, rv, zv, s);
  kb_dot<T><<<G, kB, 0, s>>>(lay.nW, rbs, lay.wPid.p, act.p, rv, zv, partial.p);
  AN_CHECK_LAUNCH();
  finalize(rhoNew);
  kb_dir<T><<<G, kB, 0, s>>>(lay.nW, rbs, lay.wPid.p, act.p, pv, zv, scal.p, rhoNew, rhoOld, 0);
  AN_CHECK_LAUNCH();
}

template <typename T>
hipGraphExec_t BatchSolver::capture(const DevCSR<T>& M, T* xv, T* rv, T* zv, T* pv, T* qv) {
  hipGraph_t graph = nullptr;
  AN_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  for (int it = 1; it <= kGraphIters; ++it) pcg_iteration<T>(M, it, xv, rv, zv, pv, qv);
  AN_HIP(hipStreamEndCapture(s, &graph));
  hipGraphExec_t exec = nullptr;
  AN_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
  (void)hipGraphDestroy(graph);
  return exec;
}

// PCG on every active problem until none is active; returns graph launches
template <typename T>
int BatchSolver::run_pcg(const DevCSR<T>& M, T* xv, T* rv, T* zv, T* pv, T* qv,
                         hipGraphExec_t& g, int maxIter) {
  pcg_prologue<T>(M, xv, rv, zv, pv);
  if (useGraph && !g) g = capture<T>(M, xv, rv, zv, pv, qv);
  graphIters = kGraphIters;
  int launches = 0;
  for (int done = 0; done < maxIter; done += kGraphIters) {
    if (useGraph)
      AN_HIP(hipGraphLaunch(g, s));
    else
      for (int it = 1; it <= kGraphIters; ++it) pcg_iteration<T>(M, it, xv, rv, zv, pv, qv);
    AN_HIP(hipMemcpyAsync(hAct, nAct.p, sizeof(int), hipMemcpyDeviceToHost, s));
    AN_HIP(hipStreamSynchronize(s));
    ++launches;
    if (hAct[0] == 0) break;
  }
  return launches;
}

__global__ void kb_zero_range(double* __restrict__ a, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0.0;
}

an_status BatchSolver::solve(const std::vector<const double*>& b, const std::vector<double*>& x,
                             bool dev, std::vector<an_solve_info>& info) {
  if (!valuesSet) throw Error(AN_EINVAL, "batch has no matrix values");
  auto t0 = std::chrono::steady_clock::now();
  const hipMemcpyKind kin = dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
  for (int pp = 0; pp < nP; ++pp) {
    const size_t m = (size_t)(pOff[pp + 1] - pOff[pp]);
    AN_HIP(hipMemcpyAsync(bB.p + pOff[pp], b[pp], sizeof(double) * m, kin, s));
    AN_HIP(hipMemcpyAsync(xB.p + pOff[pp], x[pp], sizeof(double) * m, kin, s));
  }
  const int G = grid_for(lay.nW, kW);
  kb_dot<double><<<G, kB, 0, s>>>(lay.nW, lay.rb.start.p, lay.wPid.p, nullptr, bB.p, bB.p,
                                  partial.p);
  AN_CHECK_LAUNCH();
  finalize(S_BB);
  read_scalars();
  std::vector<double> bn(nP), tolv(nP), h2(2 * (size_t)nP);
  std::vector<int> hact(nP, 0), zeros(nP, 0);
  for (int pp = 0; pp < nP; ++pp) {
    bn[pp] = std::sqrt(hScal[(size_t)S_COUNT * pp + S_BB]);
    tolv[pp] = std::fmax(opt.relTol * bn[pp], opt.absTol);
    if (bn[pp] == 0.0)  // HYPRE convention: b = 0 -> x = 0
      kb_zero_range<<<grid_for(pOff[pp + 1] - pOff[pp], kB), kB, 0, s>>>(
          xB.p + pOff[pp], pOff[pp + 1] - pOff[pp]);
  }
  residual64();
  read_scalars();
  std::vector<double> rn(nP);
  for (int pp = 0; pp < nP; ++pp) {
    rn[pp] = std::sqrt(hScal[(size_t)S_COUNT * pp + S_RR]);
    hact[pp] = (bn[pp] > 0.0 && rn[pp] > tolv[pp]) ? 1 : 0;
    h2[pp] = tolv[pp] * tolv[pp];
  }
  its.upload(zeros, s);
  std::vector<int> steps(nP, 0);
  if (opt.precision != 1) {
    tol2.upload(h2, s);
    act.upload(hact, s);
    run_pcg<double>(A, xB.p, r.p, z.p, p.p, q.p, g64, opt.maxIter);
    residual64();
    read_scalars();
    for (int pp = 0; pp < nP; ++pp) rn[pp] = std::sqrt(hScal[(size_t)S_COUNT * pp + S_RR]);
  } else {
    std::vector<int> outer = hact;
    std::vector<int> hits(nP, 0);
    for (int step = 0; step < 64; ++step) {
      bool any = false;
      for (int pp = 0; pp < nP; ++pp) any |= outer[pp] != 0;
      if (!any) break;
      // inner FP32 PCG from e = 0 on r32 = r, to innerRelTol * ||r|| (not
      // below half the outer target)
      for (int pp = 0; pp < nP; ++pp)
        h2[nP + pp] = outer[pp] ? std::pow(std::fmax(opt.innerRelTol * rn[pp], 0.5 * tolv[pp]), 2)
                                : 0.0;
      tol2.upload(h2, s);
      act.upload(outer, s);
      copy_d2f(n, r.p, r32.p, s);
      fill_zero(e32.p, sizeof(float) * (size_t)n, s);
      run_pcg<float>(amg.lv[0].A32, e32.p, r32.p, z32.p, p32.p, q32.p, g32, opt.maxIter);
      axpy_f2d(n, xB.p, e32.p, s);
      residual64();
      read_scalars();
      AN_HIP(hipMemcpyAsync(hits.data(), its.p, sizeof(int) * nP, hipMemcpyDeviceToHost, s));
      AN_HIP(hipStreamSynchronize(s));
      for (int pp = 0; pp < nP; ++pp) {
        if (!outer[pp]) continue;
        const double rnNew = std::sqrt(hScal[(size_t)S_COUNT * pp + S_RR]);
        ++steps[pp];
        const bool stall = !std::isfinite(rnNew) || (steps[pp] > 3 && rnNew > 0.9 * rn[pp]);
        rn[pp] = rnNew;
        if (rnNew <= tolv[pp] || stall || hits[pp] >= opt.maxIter) outer[pp] = 0;
      }
    }
  }
  std::vector<int> hits(nP, 0);
  AN_HIP(hipMemcpyAsync(hits.data(), its.p, sizeof(int) * nP, hipMemcpyDeviceToHost, s));
  const hipMemcpyKind kout = dev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
  for (int pp = 0; pp < nP; ++pp)
    AN_HIP(hipMemcpyAsync(x[pp], xB.p + pOff[pp],
                          sizeof(double) * (size_t)(pOff[pp + 1] - pOff[pp]), kout, s));
  AN_HIP(hipStreamSynchronize(s));
  lastSolveSeconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  info.assign(nP, an_solve_info{});
  an_status st = AN_OK;
  for (int pp = 0; pp < nP; ++pp) {
    an_solve_info& f = info[pp];
    f.iterations = hits[pp];
    f.converged = (bn[pp] == 0.0 || rn[pp] <= tolv[pp]) ? 1 : 0;
    f.finalResNorm = bn[pp] == 0.0 ? 0.0 : rn[pp];
    f.finalRelResNorm = bn[pp] == 0.0 ? 0.0 : rn[pp] / bn[pp];
    f.levels = amg.levels();
    f.refinementSteps = steps[pp];
    f.solveSeconds = lastSolveSeconds;
    if (!f.converged) st = AN_ENOCONV;
  }
  return st;
}

}  // namespace an
