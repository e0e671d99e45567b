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
// * Autotuner: the block height (64, 32, 16 or 4 rows per wave, or one wave
//   per row) is chosen by timing the batched SpMV on the actual matrix when
//   values are set, and the launch (one persistent kernel, or a replayed
//   graph of four iterations) by timing one PCG iteration of each.
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

// The batch kernels' bodies take the (virtual) block index lb and the
// block's LDS, so the per-launch kernels below and the persistent kernel
// (one launch for the whole solve) run the same arithmetic.

// MODE 0: y = A x; 1: y = b - A x.  DOT 1: dw.y, 2: y.y -> partial[wave]
template <typename T, int MODE, int DOT>
__device__ __forceinline__ void spmv_body(int lb, double (*prod)[kStreamCap], int nW,
                                          const int2* __restrict__ rbs,
                                          const int* __restrict__ wPid,
                                          const int* __restrict__ act,
                                          const int* __restrict__ rowOff,
                                          const int* __restrict__ cols,
                                          const T* __restrict__ vals, const T* __restrict__ x,
                                          const double* __restrict__ b, T* __restrict__ y,
                                          const T* __restrict__ dw,
                                          double* __restrict__ partial) {
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int gw = lb * kW + w;
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
  wave_sync();  // the wave's LDS slice is free for its next block
}

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
  spmv_body<T, MODE, DOT>(xcd_block(gridDim.x), prod, nW, rbs, wPid, act, rowOff, cols, vals, x,
                          b, y, dw, partial);
}

// rows of the wave's block: lane -> row (rows <= 64 per block)
__device__ __forceinline__ int wave_row(int nW, const int2* rbs, int gw, int lane, bool on) {
  if (!on) return -1;
  const int r0 = rbs[gw].x, r1 = rbs[gw + 1].x;
  return (r0 + lane < r1) ? r0 + lane : -1;
}

template <typename T>
__device__ __forceinline__ void dot_body(int lb, int nW, const int2* __restrict__ rbs,
                                         const int* __restrict__ wPid,
                                         const int* __restrict__ act, const T* __restrict__ a,
                                         const T* __restrict__ b, double* __restrict__ partial) {
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int gw = lb * kW + w;
  const int i = wave_row(nW, rbs, gw, lane, wave_on(gw, nW, wPid, act));
  double v = (i >= 0) ? (double)a[i] * (double)b[i] : 0.0;
  v = bwsum(v);
  if (lane == 0 && gw < nW) partial[gw] = v;
}

template <typename T>
__global__ __launch_bounds__(kB) void kb_dot(int nW, const int2* __restrict__ rbs,
                                             const int* __restrict__ wPid,
                                             const int* __restrict__ act, const T* __restrict__ a,
                                             const T* __restrict__ b,
                                             double* __restrict__ partial) {
  dot_body<T>(xcd_block(gridDim.x), nW, rbs, wPid, act, a, b, partial);
}

// x += alpha p ; r -= alpha q ; partial = r.r   (alpha = num / den per problem)
template <typename T>
__device__ __forceinline__ void update_body(int lb, int nW, const int2* __restrict__ rbs,
                                            const int* __restrict__ wPid,
                                            const int* __restrict__ act, T* __restrict__ x,
                                            T* __restrict__ r, const T* __restrict__ p,
                                            const T* __restrict__ q,
                                            const double* __restrict__ scal, int num, int den,
                                            double* __restrict__ partial) {
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int gw = lb * kW + w;
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

template <typename T>
__global__ __launch_bounds__(kB) void kb_update(int nW, const int2* __restrict__ rbs,
                                                const int* __restrict__ wPid,
                                                const int* __restrict__ act, T* __restrict__ x,
                                                T* __restrict__ r, const T* __restrict__ p,
                                                const T* __restrict__ q,
                                                const double* __restrict__ scal, int num, int den,
                                                double* __restrict__ partial) {
  update_body<T>(xcd_block(gridDim.x), nW, rbs, wPid, act, x, r, p, q, scal, num, den, partial);
}

// p = z + beta p (beta = num / den), or p = z
template <typename T>
__device__ __forceinline__ void dir_body(int lb, int nW, const int2* __restrict__ rbs,
                                         const int* __restrict__ wPid,
                                         const int* __restrict__ act, T* __restrict__ p,
                                         const T* __restrict__ z,
                                         const double* __restrict__ scal, int num, int den,
                                         int first) {
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int gw = lb * kW + w;
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

template <typename T>
__global__ __launch_bounds__(kB) void kb_dir(int nW, const int2* __restrict__ rbs,
                                             const int* __restrict__ wPid,
                                             const int* __restrict__ act, T* __restrict__ p,
                                             const T* __restrict__ z,
                                             const double* __restrict__ scal, int num, int den,
                                             int first) {
  dir_body<T>(xcd_block(gridDim.x), nW, rbs, wPid, act, p, z, scal, num, den, first);
}

// scal[p][slot] = sum of problem p's wave partials in block order
__device__ __forceinline__ void finalize_body(int p, double* sh, const int* __restrict__ wOff,
                                              const double* __restrict__ partial,
                                              double* __restrict__ scal, int slot) {
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
  __syncthreads();  // sh is reused by the block's next problem
}

__global__ __launch_bounds__(kB) void kb_finalize(const int* __restrict__ wOff,
                                                  const double* __restrict__ partial,
                                                  double* __restrict__ scal, int slot) {
  __shared__ double sh[kW];
  finalize_body(blockIdx.x, sh, wOff, partial, scal, slot);
}

// stop test after the update (pcg64's): converged, breakdown or out of
// iterations -> inactive.  nAct = number of problems still active.
__device__ __forceinline__ void check_body(int* cnt, int nP, const double* __restrict__ scal,
                                           const double* __restrict__ tol2,
                                           int* __restrict__ act, int* __restrict__ its,
                                           int* __restrict__ nAct, int maxIter) {
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

__global__ __launch_bounds__(kB) void kb_check(int nP, const double* __restrict__ scal,
                                               const double* __restrict__ tol2,
                                               int* __restrict__ act, int* __restrict__ its,
                                               int* __restrict__ nAct, int maxIter) {
  __shared__ int cnt[kW];
  check_body(cnt, nP, scal, tol2, act, its, nAct, maxIter);
}

// -------------------------------------------- one persistent launch ----
// The whole PCG loop of a batch as ONE kernel: a grid of co-resident
// workgroups walks the phases of each iteration (SpMV + p.q, per-problem
// sums, update + r.r, stop test, the V-cycle level by level, r.z,
// direction), every phase spread over the grid by virtual block index and
// separated from the next by a grid barrier, until no problem is active.
// Each phase runs the arithmetic of the matching launch of the per-kernel
// path (the bodies above, and the AMG kernels' expressions below: CSR-stream
// row sums in row order, member-order restriction, chunked dense solve in
// chunk order), so the results carry the same bits.  It is an ordinary
// launch, so it is captured in a HIP graph and replayed per solve.
//
// Barrier: every wave first waits until its own stores are acknowledged
// (s_barrier alone does not wait for them), then the block's first thread
// releases the block's writes at agent scope, adds the block to an arrival
// count (the last arrival resets the count and advances a generation word
// with release semantics) and polls the generation with relaxed loads -- an
// acquiring poll would invalidate the XCD's L2 on every spin while other
// blocks still work -- and acquires once it moved; after the block barrier
// every wave acquires too.  Every wait is bounded: after waitTicks of the
// wall clock (or when any block has given up) a block sets the abort word
// and leaves, so each block leaves the loop after at most one more phase
// and the kernel always ends; the host then reports the abort as an error.
// Arrivals are counted in kBarGroups sub-counters (blocks b with
// b mod kBarGroups == k; 128 bytes apart), whose last arrivals then count
// on the top counter: no single address takes every block's atomic.
static constexpr int kBarGroups = 8;
struct GridBar {
  unsigned count, gen;
  int abort, pad;
  unsigned pad2[28];
  unsigned sub[kBarGroups][32];
};

__device__ __forceinline__ bool grid_sync(GridBar* gb, long long waitTicks) {
  __builtin_amdgcn_s_waitcnt(0);  // this wave's loads and stores have completed
  __syncthreads();
  __shared__ int ok;
  if (threadIdx.x == 0) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    int good = 1;
    const unsigned g = __hip_atomic_load(&gb->gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned nb = gridDim.x;
    const unsigned groups = nb < (unsigned)kBarGroups ? nb : (unsigned)kBarGroups;
    const unsigned grp = blockIdx.x % groups;
    const unsigned members = (nb - grp + groups - 1) / groups;
    bool last = false;
    const unsigned a = __hip_atomic_fetch_add(&gb->sub[grp][0], 1u, __ATOMIC_RELAXED,
                                              __HIP_MEMORY_SCOPE_AGENT);
    if (a == members - 1) {  // the group's last arrival: reset it, count it above
      __hip_atomic_store(&gb->sub[grp][0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const unsigned t =
          __hip_atomic_fetch_add(&gb->count, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      last = t == groups - 1;
    }
    if (last) {
      __hip_atomic_store(&gb->count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&gb->gen, g + 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      const long long deadline = wall_clock64() + waitTicks;
      while (__hip_atomic_load(&gb->gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == g) {
        if (__hip_atomic_load(&gb->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0 ||
            wall_clock64() > deadline) {
          __hip_atomic_store(&gb->abort, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          good = 0;
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    if (__hip_atomic_load(&gb->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) good = 0;
    ok = good;
  }
  __syncthreads();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  return ok != 0;
}

// One AMG level as the persistent kernel sees it (x, r, y, b: the level's
// work vectors as the per-kernel cycle uses them; level 0's b is the PCG
// residual and its x the preconditioned residual z).
template <typename T>
struct PLevel {
  int n, nc, nrb, dense;
  const int2* rbs;
  const int *ro, *ci;
  const T *va, *omd;
  T *b, *x, *r, *y;
  const int *agg, *moff, *mem;
  const double* invT;
};

template <typename T>
struct PArgs {
  int nW, nP, maxIter, nLev, sweeps, itFirst;
  const int2* rbs;
  const int *wPid, *wOff;
  const int *ro, *ci;
  const T* va;
  T *x, *r, *z, *p, *q;
  double *scal, *partial;
  const double* tol2;
  int *act, *its, *nAct;
  double* dpart;  // chunk partials of the coarsest dense solve
  GridBar* bar;
  long long waitTicks;  // wall-clock ticks after which a barrier wait gives up
  const PLevel<T>* lv;
};

// x = from + omdinv (b - A from)  (KIND 1)  or  res = b - A from  (KIND 0),
// four row blocks per virtual block (k_level_rows without inner products)
template <typename T, int KIND>
__device__ __forceinline__ void rows_body(int lb, double (*prod)[kStreamCap], const PLevel<T>& L,
                                          const T* __restrict__ v, T* __restrict__ out) {
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const WaveRows q = wave_rows(L.nrb, L.rbs, lb * kW + w);
  const int i = q.r0 + lane;
  int kb = 0, ke = 0;
  double bi = 0.0, vi = 0.0, oi = 0.0;
  if (i < q.r1) {
    kb = L.ro[i];
    ke = L.ro[i + 1];
    bi = (double)L.b[i];
    if (KIND == 1) {
      vi = (double)v[i];
      oi = (double)L.omd[i];
    }
  }
  if (q.stream && q.k1 > q.k0)
    stage_products(q, lane, L.ci, L.va, [&](int j) { return (double)v[j]; }, prod[w]);
  wave_sync();
  if (i < q.r1) {
    double sum = 0.0;
    if (q.stream)
      for (int k = kb; k < ke; ++k) sum += prod[w][k - q.k0];
    else
      for (int k = kb; k < ke; ++k) sum += (double)L.va[k] * (double)v[L.ci[k]];
    out[i] = KIND == 0 ? (T)(bi - sum) : (T)(vi + oi * (bi - sum));
  }
  wave_sync();
}

template <typename T, int KIND>
__device__ __forceinline__ void rows_phase(double (*prod)[kStreamCap], const PLevel<T>& L,
                                           const T* v, T* out) {
  const int nvb = (L.nrb + kW - 1) / kW;
  for (int vb = blockIdx.x; vb < nvb; vb += gridDim.x) rows_body<T, KIND>(vb, prod, L, v, out);
}

template <typename T>
__device__ __forceinline__ void scale_phase(int n, const T* __restrict__ omd,
                                            const T* __restrict__ b, T* __restrict__ y) {
  for (int i = blockIdx.x * kB + threadIdx.x; i < n; i += gridDim.x * kB)
    y[i] = (T)((double)omd[i] * (double)b[i]);
}

// bc_I = sum of the members' residuals in member order, yc_I = omdc_I bc_I
template <typename T>
__device__ __forceinline__ void restrict_phase(const PLevel<T>& L, const PLevel<T>& C,
                                               const T* __restrict__ res) {
  for (int I = blockIdx.x * kB + threadIdx.x; I < L.nc; I += gridDim.x * kB) {
    double sum = 0.0;
    for (int k = L.moff[I]; k < L.moff[I + 1]; ++k) sum += (double)res[L.mem[k]];
    const T bI = (T)sum;
    C.b[I] = bI;
    C.y[I] = (T)((double)C.omd[I] * (double)bI);
  }
}

// e_i = y_i + xc[agg_i]
template <typename T>
__device__ __forceinline__ void prolong_phase(const PLevel<T>& L, const T* y,
                                              const T* __restrict__ xc, T* e) {
  for (int i = blockIdx.x * kB + threadIdx.x; i < L.n; i += gridDim.x * kB)
    e[i] = (T)((double)y[i] + (double)xc[L.agg[i]]);
}

// coarsest, dense: k_dense_solve's order -- per row, each chunk of CH
// columns summed in batches of 16, then the chunk sums in chunk order.
// Phase 1 (one wave per (64-row block, chunk)): chunk sums into dpart.
template <typename T, int CH>
__device__ __forceinline__ void dense_chunks(const PLevel<T>& L, double* __restrict__ dpart) {
  const int n = L.n, nch = (n + CH - 1) / CH, nrb = (n + 63) / 64;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int item = blockIdx.x * kW + w; item < nrb * nch; item += gridDim.x * kW) {
    const int rb = item / nch, c = item % nch;
    const int i = rb * 64 + lane, ic = min(i, n - 1);
    double t = 0.0;
    for (int j0 = c * CH; j0 < c * CH + CH; j0 += 16) {
      double a[16], bb[16];
#pragma unroll
      for (int u = 0; u < 16; ++u) {
        const int j = min(j0 + u, n - 1);
        a[u] = L.invT[(size_t)j * n + ic];
        bb[u] = (double)L.b[j];
      }
#pragma unroll
      for (int u = 0; u < 16; ++u)
        if (j0 + u < n) t += a[u] * bb[u];
    }
    if (i < n) dpart[(size_t)c * n + i] = t;
  }
}
// Phase 2: x_i = the chunk sums in chunk order
template <typename T, int CH>
__device__ __forceinline__ void dense_rows(const PLevel<T>& L, const double* __restrict__ dpart) {
  const int n = L.n, nch = (n + CH - 1) / CH;
  for (int i = blockIdx.x * kB + threadIdx.x; i < n; i += gridDim.x * kB) {
    double sum = 0.0;
    for (int c = 0; c < nch; ++c) sum += dpart[(size_t)c * n + i];
    L.x[i] = (T)sum;
  }
}

// coarsest without a dense inverse: one damped Jacobi sweep (k_jacobi)
template <typename T>
__device__ __forceinline__ void jacobi_phase(const PLevel<T>& L, const T* __restrict__ x,
                                             T* __restrict__ y) {
  for (int i = blockIdx.x * kB + threadIdx.x; i < L.n; i += gridDim.x * kB) {
    double sum = 0.0;
    for (int k = L.ro[i]; k < L.ro[i + 1]; ++k) sum += (double)L.va[k] * (double)x[L.ci[k]];
    y[i] = (T)((double)x[i] + (double)L.omd[i] * ((double)L.b[i] - sum));
  }
}

template <typename T>
__device__ __forceinline__ void fill_phase(int n, T* __restrict__ a) {
  for (int i = blockIdx.x * kB + threadIdx.x; i < n; i += gridDim.x * kB) a[i] = (T)0;
}

#define AN_SYNC()                            \
  do {                                       \
    if (!grid_sync(A.bar, A.waitTicks)) return; \
  } while (0)

// The pre-smoothed iterate of a level after sweeps - 1 sweeps that
// ping-pong from y between x and r (vcycle()'s xs): y, x, r, x, ...
template <typename T>
__device__ __forceinline__ T* presmoothed(const PLevel<T>& L, int sweeps) {
  return sweeps == 1 ? L.y : (sweeps % 2 == 0 ? L.x : L.r);
}

// The V-cycle of vcycle() (amg.hip) from x = 0 with Jacobi smoothing,
// level by level with a barrier after every phase.  Buffer choices follow
// vcycle() so each level's vectors hold what they hold there.
template <typename T>
__device__ void persist_vcycle(const PArgs<T>& A, double (*prod)[kStreamCap]) {
  const int last = A.nLev - 1;
  // down: y0 = omdinv o b0 (none when level 0 is the coarsest, as in
  // vcycle()), then per level: pre-sweeps, residual, restrict
  if (last > 0) {
    scale_phase<T>(A.lv[0].n, A.lv[0].omd, A.lv[0].b, A.lv[0].y);
    AN_SYNC();
  }
  for (int l = 0; l < last; ++l) {
    const PLevel<T>& L = A.lv[l];
    const T* xs = L.y;
    for (int k = 1; k < A.sweeps; ++k) {
      T* to = (xs == L.x) ? L.r : L.x;
      rows_phase<T, 1>(prod, L, xs, to);
      AN_SYNC();
      xs = to;
    }
    T* res = (xs == L.r) ? L.x : L.r;
    rows_phase<T, 0>(prod, L, xs, res);
    AN_SYNC();
    restrict_phase<T>(L, A.lv[l + 1], res);
    AN_SYNC();
  }
  // coarsest
  {
    const PLevel<T>& C = A.lv[last];
    if (C.dense) {
      if (C.n <= 1024) {
        dense_chunks<T, 16>(C, A.dpart);
        AN_SYNC();
        dense_rows<T, 16>(C, A.dpart);
      } else {
        dense_chunks<T, 64>(C, A.dpart);
        AN_SYNC();
        dense_rows<T, 64>(C, A.dpart);
      }
      AN_SYNC();
    } else {
      fill_phase<T>(C.n, C.x);
      AN_SYNC();
      for (int it = 0; it < 20; ++it) {
        const T* src = (it & 1) ? C.r : C.x;
        T* dst = (it & 1) ? C.x : C.r;
        jacobi_phase<T>(C, src, dst);
        AN_SYNC();
      }
    }
  }
  // up: prolongation, then the post sweeps
  for (int l = last - 1; l >= 0; --l) {
    const PLevel<T>& L = A.lv[l];
    T* e = (A.sweeps % 2 == 1) ? L.r : L.x;
    prolong_phase<T>(L, presmoothed(L, A.sweeps), A.lv[l + 1].x, e);
    AN_SYNC();
    const T* cur = e;
    for (int k = 0; k < A.sweeps; ++k) {
      T* to = (cur == L.r) ? L.x : L.r;
      rows_phase<T, 1>(prod, L, cur, to);
      AN_SYNC();
      cur = to;
    }
  }
}

// PCG iterations itFirst, itFirst + 1, ... until no problem is active
// (pcg_iteration's sequence, phase by phase)
template <typename T>
__global__ __launch_bounds__(kB, 4) void kb_persist(const PArgs<T>* __restrict__ pa) {
  __shared__ double prod[kW][kStreamCap];
  __shared__ double sh[kW];
  __shared__ int cnt[kW];
  const PArgs<T> A = *pa;
  const int nvbW = (A.nW + kW - 1) / kW;
  for (int it = A.itFirst; it < A.itFirst + A.maxIter + 1; ++it) {
    const int rhoOld = (it & 1) ? S_RHO0 : S_RHO1;
    const int rhoNew = (it & 1) ? S_RHO1 : S_RHO0;
    for (int vb = blockIdx.x; vb < nvbW; vb += gridDim.x)
      spmv_body<T, 0, 1>(vb, prod, A.nW, A.rbs, A.wPid, A.act, A.ro, A.ci, A.va, A.p, nullptr,
                         A.q, A.p, A.partial);
    AN_SYNC();
    for (int p = blockIdx.x; p < A.nP; p += gridDim.x)
      finalize_body(p, sh, A.wOff, A.partial, A.scal, S_PQ);
    AN_SYNC();
    for (int vb = blockIdx.x; vb < nvbW; vb += gridDim.x)
      update_body<T>(vb, A.nW, A.rbs, A.wPid, A.act, A.x, A.r, A.p, A.q, A.scal, rhoOld, S_PQ,
                     A.partial);
    AN_SYNC();
    for (int p = blockIdx.x; p < A.nP; p += gridDim.x)
      finalize_body(p, sh, A.wOff, A.partial, A.scal, S_RR);
    AN_SYNC();
    if (blockIdx.x == 0) check_body(cnt, A.nP, A.scal, A.tol2, A.act, A.its, A.nAct, A.maxIter);
    AN_SYNC();
    if (__hip_atomic_load(A.nAct, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) return;
    persist_vcycle<T>(A, prod);
    if (__hip_atomic_load(&A.bar->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0)
      return;
    for (int vb = blockIdx.x; vb < nvbW; vb += gridDim.x)
      dot_body<T>(vb, A.nW, A.rbs, A.wPid, A.act, A.r, A.z, A.partial);
    AN_SYNC();
    for (int p = blockIdx.x; p < A.nP; p += gridDim.x)
      finalize_body(p, sh, A.wOff, A.partial, A.scal, rhoNew);
    AN_SYNC();
    for (int vb = blockIdx.x; vb < nvbW; vb += gridDim.x)
      dir_body<T>(vb, A.nW, A.rbs, A.wPid, A.act, A.p, A.z, A.scal, rhoNew, rhoOld, 0);
    AN_SYNC();
  }
}
#undef AN_SYNC

__global__ void kb_bar_reset(GridBar* gb) {
  unsigned* w = reinterpret_cast<unsigned*>(gb);
  for (int i = threadIdx.x; i < (int)(sizeof(GridBar) / sizeof(unsigned)); i += blockDim.x)
    w[i] = 0u;
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
  AN_HIP(hipHostMalloc((void**)&hAct, sizeof(int) * (size_t)(nP + 4)));
  AN_HIP(hipHostMalloc((void**)&hScal, sizeof(double) * (size_t)S_COUNT * nP));
  int device = 0, rateKHz = 0;
  AN_HIP(hipGetDevice(&device));
  AN_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
  AN_HIP(hipDeviceGetAttribute(&rateKHz, hipDeviceAttributeWallClockRate, device));
  waitTicks = 2LL * 1000LL * (long long)(rateKHz > 0 ? rateKHz : 100000);  // 2 s per wait
  AN_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&maxPerCU64, kb_persist<double>, kB, 0));
  AN_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&maxPerCU32, kb_persist<float>, kB, 0));
  bar.alloc(sizeof(GridBar) / sizeof(unsigned));
  AN_HIP(hipStreamSynchronize(s));
}

BatchSolver::~BatchSolver() {
  if (g64) (void)hipGraphExecDestroy(g64);
  if (g32) (void)hipGraphExecDestroy(g32);
  gp64.drop();
  gp32.drop();
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
  // 64, 32, 16 or 4 rows per wave, or one (a wave per row: its lanes load
  // the row's nonzeros together, one lane adds them in order)
  tuneRows = {64, 32, 16, 4, 1};
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
  for (hipGraphExec_t* g : {&g64, &g32}) {
    if (*g) (void)hipGraphExecDestroy(*g);
    *g = nullptr;
  }
  gp64.drop();
  gp32.drop();
  // the coarsest level's dense solve keeps its chunk sums in dpart
  const AmgLevel& Lc = amg.lv.back();
  const int ch = Lc.n <= 1024 ? 16 : 64;
  dpart.alloc((size_t)((Lc.n + ch - 1) / ch) * (size_t)(Lc.n > 0 ? Lc.n : 1));
  valuesSet = true;
  tune_launch();
}

int BatchSolver::mode() const { return launch == AN_BATCH_AUTO ? chosen : launch; }

void BatchSolver::PersistGraph::drop() {
  if (exec) (void)hipGraphExecDestroy(exec);
  exec = nullptr;
  grid = maxIter = 0;
}

// The persistent kernel's arguments and level views, uploaded to device
// memory (the captured graph refers to them there).
template <typename T>
void BatchSolver::build_persist(const DevCSR<T>& M, T* xv, T* rv, T* zv, T* pv, T* qv,
                                int tolOff, int maxIter, PersistGraph& g) {
  DevBuf<char>& pa = g.args;
  DevBuf<char>& plv = g.levels;
  using D = LevelData<T>;
  std::vector<PLevel<T>> lv(amg.lv.size());
  for (size_t l = 0; l < amg.lv.size(); ++l) {
    AmgLevel& L = amg.lv[l];
    const DevCSR<T>& A_ = D::A(L);
    PLevel<T>& v = lv[l];
    v.n = L.n;
    v.nc = L.nc;
    v.nrb = A_.rb.count;
    v.dense = (L.isCoarsest && L.coarseInvT.p) ? 1 : 0;
    v.rbs = A_.rb.start.p;
    v.ro = A_.rowOff.p;
    v.ci = A_.cols.p;
    v.va = A_.vals.p;
    v.omd = D::omd(L);
    v.b = l == 0 ? rv : D::b(L);
    v.x = l == 0 ? zv : D::x(L);
    v.r = D::r(L);
    v.y = D::y(L);
    v.agg = L.agg.p;
    v.moff = L.moff.p;
    v.mem = L.mem.p;
    v.invT = L.coarseInvT.p;
  }
  plv.alloc(sizeof(PLevel<T>) * lv.size());
  AN_HIP(hipMemcpyAsync(plv.p, lv.data(), sizeof(PLevel<T>) * lv.size(), hipMemcpyHostToDevice,
                        s));
  PArgs<T> a;
  a.nW = lay.nW;
  a.nP = nP;
  a.maxIter = maxIter;
  a.nLev = (int)amg.lv.size();
  a.sweeps = std::max(amg.opt.sweeps, 1);
  a.itFirst = 1;
  a.rbs = lay.rb.start.p;
  a.wPid = lay.wPid.p;
  a.wOff = lay.wOff.p;
  a.ro = M.rowOff.p;
  a.ci = M.cols.p;
  a.va = M.vals.p;
  a.x = xv;
  a.r = rv;
  a.z = zv;
  a.p = pv;
  a.q = qv;
  a.scal = scal.p;
  a.partial = partial.p;
  a.tol2 = tol2.p + tolOff;
  a.act = act.p;
  a.its = its.p;
  a.nAct = nAct.p;
  a.dpart = dpart.p;
  a.bar = reinterpret_cast<GridBar*>(bar.p);
  a.waitTicks = waitTicks;
  a.lv = reinterpret_cast<const PLevel<T>*>(plv.p);
  pa.alloc(sizeof(PArgs<T>));
  AN_HIP(hipMemcpyAsync(pa.p, &a, sizeof(PArgs<T>), hipMemcpyHostToDevice, s));
  AN_HIP(hipStreamSynchronize(s));  // the host copies of a and lv go out of scope
}

// The PCG loop as one persistent launch: {reset the barrier words, the
// kernel} captured once per grid and iteration cap, then replayed.
template <typename T>
void BatchSolver::launch_persist(const DevCSR<T>& M, T* xv, T* rv, T* zv, T* pv, T* qv,
                                 int tolOff, PersistGraph& g, int blocksPerCU, int maxIter) {
  const int maxPer = sizeof(T) == 8 ? maxPerCU64 : maxPerCU32;
  const int per = std::min(blocksPerCU, maxPer);
  if (per < 1) throw Error(AN_EHIP, "the persistent batch kernel does not fit on a CU");
  const int grid = per * cus;
  if (!g.exec || g.grid != grid || g.maxIter != maxIter) {
    g.drop();
    build_persist<T>(M, xv, rv, zv, pv, qv, tolOff, maxIter, g);
    hipGraph_t graph = nullptr;
    AN_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    kb_bar_reset<<<1, 64, 0, s>>>(reinterpret_cast<GridBar*>(bar.p));
    kb_persist<T><<<grid, kB, 0, s>>>(reinterpret_cast<const PArgs<T>*>(g.args.p));
    AN_HIP(hipStreamEndCapture(s, &graph));
    const hipError_t e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    AN_HIP(e);
    g.grid = grid;
    g.maxIter = maxIter;
  }
  AN_HIP(hipGraphLaunch(g.exec, s));
}

// the abort word says whether every barrier of the last launch completed
void BatchSolver::check_persist() {
  AN_HIP(hipMemcpyAsync(hAct, bar.p, sizeof(unsigned) * 4, hipMemcpyDeviceToHost, s));
  AN_HIP(hipStreamSynchronize(s));
  if (hAct[2] != 0)
    throw Error(AN_EHIP,
                "persistent batch solve: a grid barrier wait timed out (workgroups not resident)");
}

__global__ void kb_fill(int n, double* __restrict__ a, double v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

// The launch tuner: one PCG iteration's time (4 iterations from r = 1, every
// problem active) by graph replay and by the persistent kernel at 1, 2 and 4
// workgroups per CU; the fastest becomes the AN_BATCH_AUTO launch.
void BatchSolver::tune_launch() {
  constexpr int kIters = kGraphIters;
  tuneBlocksPerCU = {1, 2, 4};
  launchSeconds.assign(1 + tuneBlocksPerCU.size(), 0.0);
  std::vector<int> ones(nP, 1), zeros(nP, 0);
  std::vector<double> tz(2 * (size_t)nP, 0.0);
  hipEvent_t e0, e1;
  AN_HIP(hipEventCreate(&e0));
  AN_HIP(hipEventCreate(&e1));
  auto timed = [&](int which) {
    double best = 1e300;
    for (int rep = 0; rep < 2; ++rep) {  // the first run also captures the graph
      kb_fill<<<grid_for(n, kB), kB, 0, s>>>(n, r.p, 1.0);
      fill_zero(xB.p, sizeof(double) * (size_t)n, s);
      act.upload(ones, s);
      its.upload(zeros, s);
      tol2.upload(tz, s);
      pcg_prologue<double>(A, xB.p, r.p, z.p, p.p);
      if (which == 0 && !g64) g64 = capture<double>(A, xB.p, r.p, z.p, p.p, q.p);
      AN_HIP(hipEventRecord(e0, s));
      if (which == 0)
        AN_HIP(hipGraphLaunch(g64, s));
      else
        launch_persist<double>(A, xB.p, r.p, z.p, p.p, q.p, 0, gp64,
                               tuneBlocksPerCU[which - 1], kIters);
      AN_HIP(hipEventRecord(e1, s));
      AN_HIP(hipEventSynchronize(e1));
      if (which > 0) check_persist();
      float ms = 0.f;
      AN_HIP(hipEventElapsedTime(&ms, e0, e1));
      if (rep > 0) best = std::min(best, ms * 1e-3 / kIters);
    }
    return best;
  };
  double best = 1e300;
  chosen = AN_BATCH_GRAPH;
  persistBlocksPerCU = 1;
  for (size_t c = 0; c < launchSeconds.size(); ++c) {
    // a grid above the resident maximum is clipped to it (launch_persist)
    if (c > 0 && std::min(maxPerCU64, maxPerCU32) < 1) continue;
    launchSeconds[c] = timed((int)c);
    if (launchSeconds[c] < best) {
      best = launchSeconds[c];
      chosen = c == 0 ? AN_BATCH_GRAPH : AN_BATCH_PERSISTENT;
      if (c > 0)
        persistBlocksPerCU =
            std::min(tuneBlocksPerCU[c - 1], std::min(maxPerCU64, maxPerCU32));
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  gp64.drop();  // solves capture their own graph (their iteration cap)
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

// PCG on every active problem until none is active; returns the number of
// kernel-sequence launches (1 for the persistent kernel)
template <typename T>
int BatchSolver::run_pcg(const DevCSR<T>& M, T* xv, T* rv, T* zv, T* pv, T* qv,
                         hipGraphExec_t& g, PersistGraph& gp, int maxIter) {
  pcg_prologue<T>(M, xv, rv, zv, pv);
  const int md = mode();
  if (md == AN_BATCH_PERSISTENT) {
    launch_persist<T>(M, xv, rv, zv, pv, qv, sizeof(T) == 4 ? nP : 0, gp, persistBlocksPerCU,
                      maxIter);
    check_persist();
    graphIters = 0;
    return 1;
  }
  const bool graph = md == AN_BATCH_GRAPH;
  if (graph && !g) g = capture<T>(M, xv, rv, zv, pv, qv);
  graphIters = kGraphIters;
  int launches = 0;
  for (int done = 0; done < maxIter; done += kGraphIters) {
    if (graph)
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
    lastLaunches = run_pcg<double>(A, xB.p, r.p, z.p, p.p, q.p, g64, gp64, opt.maxIter);
    residual64();
    read_scalars();
    for (int pp = 0; pp < nP; ++pp) rn[pp] = std::sqrt(hScal[(size_t)S_COUNT * pp + S_RR]);
  } else {
    lastLaunches = 0;
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
      lastLaunches += run_pcg<float>(amg.lv[0].A32, e32.p, r32.p, z32.p, p32.p, q32.p, g32,
                                     gp32, opt.maxIter);
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
