// linalg.hip -- CSR SpMV, fused Krylov vector updates and deterministic
// reductions for the PCG/BiCGSTAB solvers.
//
// SpMV (CSR-stream, wave granularity, after Greathouse & Daga's CSR-Adaptive,
// SC'14): each 64-lane wavefront owns one row block (at most 64 consecutive
// rows holding at most 512 nonzeros, built once per pattern).  The lanes
// first stream the block's nonzeros with fully coalesced loads of vals/cols
// (lane k reads nonzero k0+k), multiply by the gathered x and stage the
// products in the wave's LDS slice; then lane i sums row i's products from
// LDS in CSR order.  Rows are 4-27 entries long on polyhedral meshes and
// 20-60 on AMG coarse levels, so thread-per-row loads would be strided by
// the row length; the stream form turns them into unit-stride loads.  A row
// longer than the LDS slice forms its own block and is summed by one lane
// from global memory.  Sums run in CSR order, so y equals the serial product
// bit for bit.
//
// Reductions: every block writes a partial sum (fixed element->block map,
// fixed shuffle tree) and a one-block finalize kernel adds the partials in
// index order into a device scalar: results are deterministic, no global
// atomics are involved, and scalars never leave the GPU inside an iteration.
#include "linalg.hpp"

#include <hip/hip_ext.h>

namespace an {

static constexpr int kBlock = 256;
static constexpr int kWaves = kBlock / 64;
static constexpr int kCap = kStreamCap;  // staged products per wave (4 KB)
static constexpr int kRedBlocks = 1024;

// ------------------------------------------------------------ row blocks
// Fixed 64-row slices are split greedily where their nonzeros exceed the
// cap; pass 0 counts the blocks of each slice, pass 1 writes their starts.
template <int PASS>
__global__ void k_row_blocks(int n, const int* __restrict__ ro, const int* __restrict__ off,
                             int* __restrict__ cnto, int2* __restrict__ out) {
  const int sl = blockIdx.x * blockDim.x + threadIdx.x;
  const int nsl = (n + kStreamRows - 1) / kStreamRows;
  if (sl >= nsl) return;
  const int a = sl * kStreamRows, b = min(a + kStreamRows, n);
  int cnt = 0, r = a;
  while (r < b) {
    const int st = r;
    ++r;
    while (r < b && ro[r + 1] - ro[st] <= kStreamCap) ++r;
    if (PASS == 1) out[off[sl] + cnt] = make_int2(st, ro[st]);
    ++cnt;
  }
  if (PASS == 0) cnto[sl] = cnt;
  if (PASS == 1 && sl == nsl - 1) out[off[sl] + cnt] = make_int2(n, ro[n]);
}

void build_row_blocks(int n, const int* rowOff, RowBlocks& rb, hipStream_t s) {
  const int nsl = (n + kStreamRows - 1) / kStreamRows;
  DevBuf<int> cnt((size_t)nsl + 1), off((size_t)nsl + 1);
  rb.count = 0;
  if (nsl == 0) return;
  k_row_blocks<0><<<grid_for(nsl, 256), 256, 0, s>>>(n, rowOff, nullptr, cnt.p, nullptr);
  AN_CHECK_LAUNCH();
  const long long tot = exclusive_scan(cnt.p, off.p, nsl, s);
  rb.count = (int)tot;
  rb.start.alloc((size_t)tot + 1);
  k_row_blocks<1><<<grid_for(nsl, 256), 256, 0, s>>>(n, rowOff, off.p, nullptr, rb.start.p);
  AN_CHECK_LAUNCH();
  AN_HIP(hipStreamSynchronize(s));
}

// one thread per row block: its nonzeros' columns relative to its first row
// one wave per row block, its lanes across the block's nonzeros
__global__ __launch_bounds__(256) void k_col_offsets(int nrb, const int2* __restrict__ rbs,
                                                     const int* __restrict__ cols,
                                                     int16_t* __restrict__ c16,
                                                     int* __restrict__ overflow) {
  const int b = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (b >= nrb) return;  // wave-uniform
  const int2 a = rbs[b], e = rbs[b + 1];
  bool over = false;
  for (int k = a.y + lane; k < e.y; k += 64) {
    const int d = cols[k] - a.x;
    over |= d < -32768 || d > 32767;
    c16[k] = (int16_t)d;  // not used when any offset overflows
  }
  if (over) atomicOr(overflow, 1);
}

bool build_col_offsets(long long nnz, const int* cols, const RowBlocks& rb, DevBuf<int16_t>& c16,
                       hipStream_t s) {
  c16.release();
  if (nnz <= 0 || rb.count == 0) return false;
  c16.alloc((size_t)nnz);
  DevBuf<int> over(1);
  over.zero(s);
  k_col_offsets<<<grid_for(rb.count, 4), 256, 0, s>>>(rb.count, rb.start.p, cols, c16.p, over.p);
  AN_CHECK_LAUNCH();
  int h = 0;
  AN_HIP(hipMemcpyAsync(&h, over.p, sizeof(int), hipMemcpyDeviceToHost, s));
  AN_HIP(hipStreamSynchronize(s));
  if (h) c16.release();
  return h == 0;
}

__device__ __forceinline__ double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// sum over the block (valid in thread 0)
__device__ __forceinline__ double block_sum(double v, double* sh) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) t += sh[i];
  return t;
}

// Two-stage deterministic reduction: every block stores its partial, then
// k_finalize (one block) adds the partials in index order.

__global__ __launch_bounds__(1024) void k_finalize(int nparts, int comps,
                                                   const double* __restrict__ partial,
                                                   double* __restrict__ out) {
  __shared__ double sh[16];
  for (int c = 0; c < comps; ++c) {
    double s = 0.0;
    for (int i = threadIdx.x; i < nparts; i += blockDim.x) s += partial[(size_t)i * comps + c];
    s = block_sum(s, sh);
    if (threadIdx.x == 0) out[c] = s;
    __syncthreads();
  }
}

// The same sums for a fixed component count in one pass: every thread loads
// all NC components of its partials (four partials in flight) and the NC
// block sums share one barrier round.  Each component is added in the order
// of k_finalize, so the results are the same bits.
template <int NC>
__global__ __launch_bounds__(1024) void k_finalize_nc(int nparts,
                                                      const double* __restrict__ partial,
                                                      double* __restrict__ out) {
  __shared__ double sh[NC * 16];
  double v[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) v[c] = 0.0;
#pragma unroll 4
  for (int i = threadIdx.x; i < nparts; i += blockDim.x)
#pragma unroll
    for (int c = 0; c < NC; ++c) v[c] += partial[(size_t)i * NC + c];
  block_sums<NC>(v, sh);
  if (threadIdx.x == 0)
#pragma unroll
    for (int c = 0; c < NC; ++c) out[c] = v[c];
}

static void launch_finalize(int nparts, int comps, const double* partial, double* out,
                            hipStream_t s) {
  switch (comps) {
    case 1: k_finalize_nc<1><<<1, 1024, 0, s>>>(nparts, partial, out); break;
    case 2: k_finalize_nc<2><<<1, 1024, 0, s>>>(nparts, partial, out); break;
    case 3: k_finalize_nc<3><<<1, 1024, 0, s>>>(nparts, partial, out); break;
    case 6: k_finalize_nc<6><<<1, 1024, 0, s>>>(nparts, partial, out); break;
    default: k_finalize<<<1, 1024, 0, s>>>(nparts, comps, partial, out);
  }
  AN_CHECK_LAUNCH();
}

void finalize(const Reducer& red, int nparts, int comps, double* out, hipStream_t s) {
  launch_finalize(nparts, comps, red.partial.p, out, s);
}

void finalize_parts(const RedArgs& ra, int nparts, int comps, hipStream_t s) {
  launch_finalize(nparts, comps, ra.partial, ra.out, s);
}

void Reducer::init(int maxBlocks, hipStream_t s) {
  cap = maxBlocks;
  partial.alloc(4 * (size_t)maxBlocks);
  (void)s;
}

// ------------------------------------------------------------------ SpMV
// MODE 0: y = A x ; MODE 1: y = b - A x.  DOT 0: none; 1: dot(dw, y);
// 2: dot(y, y); 3 (MODE 1): dot(y, y) and dot(b, b); 4: dot(dw, y) and
// dot(y, y) (two partials per block for 3 and 4).
// C16: columns from the 16-bit offsets c16 (ColOffsets) instead of cols.
template <typename VT, typename XT, typename YT, int MODE, int DOT, bool C16>
__global__ __launch_bounds__(kBlock) void k_spmv(int nrb, const int2* __restrict__ rbs,
                                                 const int* __restrict__ rowOff,
                                                 const int* __restrict__ cols,
                                                 const int16_t* __restrict__ c16,
                                                 const VT* __restrict__ vals,
                                                 const XT* __restrict__ x,
                                                 const double* __restrict__ b, YT* __restrict__ y,
                                                 const YT* __restrict__ dw, RedArgs ra) {
  __shared__ double prod[kWaves][kCap];
  __shared__ double sh[2 * kWaves];
  const int lb = xcd_block(gridDim.x);
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const WaveRows q = wave_rows(nrb, rbs, lb * kWaves + w);
  const int row = q.r0 + lane;
  // the row's own operands are loaded with the row offsets, so their latency
  // hides behind the nonzero stream
  int kb = 0, ke = 0;
  double bi = 0.0, wi = 0.0;
  if (row < q.r1) {
    kb = rowOff[row];
    ke = rowOff[row + 1];
    if (MODE == 1) bi = b[row];
    if (DOT == 1 || DOT == 4) wi = (double)dw[row];
  }
  auto xv = [&](int j) { return (double)x[j]; };
  if (q.stream && q.k1 > q.k0) {
    if constexpr (C16)
      stage_products(q, lane, ColOffsets{c16, q.r0}, vals, xv, prod[w]);
    else
      stage_products(q, lane, cols, vals, xv, prod[w]);
  }
  wave_sync();
  double dv = 0.0, db = 0.0;
  if (row < q.r1) {
    double s = 0.0;
    if (q.stream)
      for (int k = kb; k < ke; ++k) s += prod[w][k - q.k0];
    else
      for (int k = kb; k < ke; ++k)
        s += (double)vals[k] * (double)x[C16 ? q.r0 + (int)c16[k] : cols[k]];
    YT v;
    if (MODE == 1)
      v = (YT)(bi - s);
    else
      v = (YT)s;
    y[row] = v;
    if (DOT == 1 || DOT == 4) dv = wi * (double)v;
    if (DOT == 2 || DOT == 3) dv = (double)v * (double)v;
    if (DOT == 3) db = bi * bi;
    if (DOT == 4) db = (double)v * (double)v;
  }
  if (DOT >= 3) {
    double d2[2] = {dv, db};
    block_sums<2>(d2, sh);
    store_partials<2>(d2, lb, ra);
  } else if (DOT) {
    const double bs = block_sum(dv, sh);
    store_partials<1>(&bs, lb, ra);
  }
}

template <typename VT>
static int spmv_grid(const DevCSR<VT>& A) {
  if (A.n > 0 && A.rb.count == 0) throw Error(7, "CSR row blocks not built");
  return grid_for(A.rb.count, kWaves);
}

// k_spmv with 16-bit column offsets when the matrix has them; with
// evStart/evStop the launch is timed (hipExtLaunchKernel)
template <typename VT, typename XT, typename YT, int MODE, int DOT>
static void launch_spmv(const DevCSR<VT>& A, int nb, hipStream_t s, const XT* x, const double* b,
                        YT* y, const YT* dw, RedArgs ra, hipEvent_t evStart = nullptr,
                        hipEvent_t evStop = nullptr) {
  const int2* rbs = A.rb.start.p;
  const int* ro = A.rowOff.p;
  const int* ci = A.cols.p;
  const int16_t* c16 = A.c16.p;
  const VT* va = A.vals.p;
  if (A.has16) {
    if (evStart)
      hipExtLaunchKernelGGL(k_spmv<VT, XT, YT, MODE, DOT, true>, dim3(nb), dim3(kBlock), 0, s,
                            evStart, evStop, 0, A.rb.count, rbs, ro, ci, c16, va, x, b, y, dw, ra);
    else
      k_spmv<VT, XT, YT, MODE, DOT, true><<<nb, kBlock, 0, s>>>(A.rb.count, rbs, ro, ci, c16, va,
                                                                x, b, y, dw, ra);
  } else {
    if (evStart)
      hipExtLaunchKernelGGL(k_spmv<VT, XT, YT, MODE, DOT, false>, dim3(nb), dim3(kBlock), 0, s,
                            evStart, evStop, 0, A.rb.count, rbs, ro, ci, c16, va, x, b, y, dw, ra);
    else
      k_spmv<VT, XT, YT, MODE, DOT, false><<<nb, kBlock, 0, s>>>(A.rb.count, rbs, ro, ci, c16,
                                                                 va, x, b, y, dw, ra);
  }
  AN_CHECK_LAUNCH();
}

template <typename VT, typename XT>
void spmv(const DevCSR<VT>& A, const XT* x, XT* y, const XT* dotWith, double* scal, int slot,
          Reducer& red, hipStream_t s, hipEvent_t evStart, hipEvent_t evStop) {
  const int nb = spmv_grid(A);
  const double* nob = nullptr;
  if (dotWith) {
    if (nb > red.cap) throw Error(7, "reducer too small");
    const RedArgs ra = red.args(scal + slot);
    launch_spmv<VT, XT, XT, 0, 1>(A, nb, s, x, nob, y, dotWith, ra, evStart, evStop);
    finalize_parts(ra, nb, 1, s);
  } else {
    const XT* nodw = nullptr;
    launch_spmv<VT, XT, XT, 0, 0>(A, nb, s, x, nob, y, nodw, RedArgs{nullptr, nullptr}, evStart,
                                  evStop);
  }
}

template <typename VT, typename RT>
void residual(const DevCSR<VT>& A, const double* b, const double* x, RT* r, double* scal,
              Reducer& red, hipStream_t s) {
  const int nb = spmv_grid(A);
  if (nb > red.cap) throw Error(7, "reducer too small");
  launch_spmv<VT, double, RT, 1, 2>(A, nb, s, x, b, r, (const RT*)nullptr, red.args(scal + S_RR));
  finalize(red, nb, 1, scal + S_RR, s);
}

template <typename VT, typename RT>
int residual_bb(const DevCSR<VT>& A, const double* b, const double* x, RT* r, Reducer& red,
                hipStream_t s) {
  const int nb = spmv_grid(A);
  if (2 * nb > 4 * red.cap) throw Error(7, "reducer too small");
  launch_spmv<VT, double, RT, 1, 3>(A, nb, s, x, b, r, (const RT*)nullptr, red.args(nullptr));
  return nb;
}

int spmv_dot_self(const DevCSR<double>& A, const double* x, double* y, const double* dw,
                  Reducer& red, hipStream_t s) {
  const int nb = spmv_grid(A);
  if (2 * nb > 4 * red.cap) throw Error(7, "reducer too small");
  launch_spmv<double, double, double, 0, 4>(A, nb, s, x, (const double*)nullptr, y, dw,
                                            red.args(nullptr));
  return nb;
}

template <typename VT, typename XT>
int spmv_self_dot(const DevCSR<VT>& A, const XT* x, XT* y, Reducer& red, hipStream_t s) {
  const int nb = spmv_grid(A);
  if (nb > red.cap) throw Error(7, "reducer too small");
  launch_spmv<VT, XT, XT, 0, 2>(A, nb, s, x, (const double*)nullptr, y, (const XT*)nullptr,
                                red.args(nullptr));
  return nb;
}

// ---------------------------------------------------------- vector ops
static inline int red_blocks(int n) {
  int nb = grid_for(n, kBlock);
  return nb < kRedBlocks ? nb : kRedBlocks;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_dot(int n, const T* __restrict__ a,
                                                const T* __restrict__ b, RedArgs ra) {
  __shared__ double sh[kWaves];
  double v = 0.0;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
    v += (double)a[i] * (double)b[i];
  const double bs = block_sum(v, sh);
  store_partials<1>(&bs, blockIdx.x, ra);
}

template <typename T>
void dot(int n, const T* a, const T* b, double* scal, int slot, Reducer& red, hipStream_t s) {
  const int nb = red_blocks(n);
  k_dot<T><<<nb, kBlock, 0, s>>>(n, a, b, red.args(scal + slot));
  AN_CHECK_LAUNCH();
  finalize(red, nb, 1, scal + slot, s);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_dot2(int n, const T* __restrict__ a,
                                                 const T* __restrict__ b,
                                                 const T* __restrict__ c, RedArgs ra) {
  __shared__ double sh[2 * kWaves];
  double v[2] = {0.0, 0.0};
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    const double bi = (double)b[i];
    v[0] += (double)a[i] * bi;
    v[1] += bi * (double)c[i];
  }
  block_sums<2>(v, sh);
  store_partials<2>(v, blockIdx.x, ra);
}

template <typename T>
void dot2(int n, const T* a, const T* b, const T* c, double* scal, int slot, Reducer& red,
          hipStream_t s) {
  const int nb = red_blocks(n);
  k_dot2<T><<<nb, kBlock, 0, s>>>(n, a, b, c, red.args(scal + slot));
  AN_CHECK_LAUNCH();
  finalize(red, nb, 2, scal + slot, s);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_pcg_update(int n, T* __restrict__ x,
                                                       T* __restrict__ r,
                                                       const T* __restrict__ p,
                                                       const T* __restrict__ q,
                                                       const double* scal, int num, int den,
                                                       RedArgs ra, const T* __restrict__ omd,
                                                       T* __restrict__ y) {
  __shared__ double sh[kWaves];
  const double d = scal[den], a = scal[num] / d;
  const double alpha = (d > 0.0 && isfinite(a)) ? a : 0.0;
  double v = 0.0;
  // a thread's grid-stride elements four at a time: all loads of the four
  // first, then the updates and r.r in element order (the bits of one
  // element at a time)
  constexpr int U = 4;
  const int S = gridDim.x * kBlock;
  for (int i0 = blockIdx.x * kBlock + threadIdx.x; i0 < n; i0 += U * S) {
    T xv[U], pv[U], rv[U], qv[U], ov[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = min(i0 + u * S, n - 1);
      xv[u] = x[i];
      pv[u] = p[i];
      rv[u] = r[i];
      qv[u] = q[i];
      ov[u] = y ? omd[i] : (T)0;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = i0 + u * S;
      if (i >= n) break;
      x[i] = (T)((double)xv[u] + alpha * (double)pv[u]);
      const T ri = (T)((double)rv[u] - alpha * (double)qv[u]);
      r[i] = ri;
      if (y) y[i] = (T)((double)ov[u] * (double)ri);  // the preconditioner's first sweep
      v += (double)ri * (double)ri;
    }
  }
  const double bs = block_sum(v, sh);
  store_partials<1>(&bs, blockIdx.x, ra);
}

// k_finalize_nc<1> for scal[S_RR] that also decides the iteration's outcome
// and publishes it (PcgPub, pcg_word)
__global__ __launch_bounds__(1024) void k_finalize_pub(int nparts,
                                                       const double* __restrict__ partial,
                                                       double* __restrict__ scal, PcgPub pub) {
  __shared__ double sh[16];
  double v[1] = {0.0};
#pragma unroll 4
  for (int i = threadIdx.x; i < nparts; i += blockDim.x) v[0] += partial[i];
  block_sums<1>(v, sh);
  if (threadIdx.x == 0) {
    scal[S_RR] = v[0];
    const double rn = sqrt(v[0]), pq = scal[S_PQ];
    const int code = (!(pq > 0.0) || !isfinite(rn)) ? PCG_BREAKDOWN
                     : (rn <= pub.tol)             ? PCG_CONVERGED
                                                   : PCG_CONTINUE;
    if (pub.hist) pub.hist[pub.idx] = rn;
    // release at system scope: the history and scalars written above are
    // visible before the word that announces them
    __hip_atomic_store(pub.word, pcg_word(pub.stamp, code, (float)rn), __ATOMIC_RELEASE,
                       __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

// k_pcg_update<double> for an FP32 preconditioner: the FP64 x and r, and
// the cycle's input r32 = (float) r with its first sweep y32 = omd32 o r32
__global__ __launch_bounds__(kBlock) void k_pcg_update_p2(int n, double* __restrict__ x,
                                                          double* __restrict__ r,
                                                          const double* __restrict__ p,
                                                          const double* __restrict__ q,
                                                          const double* scal, int num, int den,
                                                          RedArgs ra,
                                                          const float* __restrict__ omd32,
                                                          float* __restrict__ r32,
                                                          float* __restrict__ y32) {
  __shared__ double sh[kWaves];
  const double d = scal[den], a = scal[num] / d;
  const double alpha = (d > 0.0 && isfinite(a)) ? a : 0.0;
  double v = 0.0;
  // four grid-stride elements per step, loads first (k_pcg_update)
  constexpr int U = 4;
  const int S = gridDim.x * kBlock;
  for (int i0 = blockIdx.x * kBlock + threadIdx.x; i0 < n; i0 += U * S) {
    double xv[U], pv[U], rv[U], qv[U];
    float ov[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = min(i0 + u * S, n - 1);
      xv[u] = x[i];
      pv[u] = p[i];
      rv[u] = r[i];
      qv[u] = q[i];
      ov[u] = omd32[i];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = i0 + u * S;
      if (i >= n) break;
      x[i] = xv[u] + alpha * pv[u];
      const double ri = rv[u] - alpha * qv[u];
      r[i] = ri;
      const float rf = (float)ri;
      r32[i] = rf;
      y32[i] = (float)((double)ov[u] * (double)rf);
      v += ri * ri;
    }
  }
  const double bs = block_sum(v, sh);
  store_partials<1>(&bs, blockIdx.x, ra);
}

void pcg_update_p2(int n, double* x, double* r, const double* p, const double* q, double* scal,
                   int num, int den, Reducer& red, hipStream_t s, const float* omd32, float* r32,
                   float* y32, const PcgPub* pub) {
  const int nb = red_blocks(n);
  k_pcg_update_p2<<<nb, kBlock, 0, s>>>(n, x, r, p, q, scal, num, den, red.args(scal + S_RR), omd32,
                                        r32, y32);
  AN_CHECK_LAUNCH();
  if (pub && pub->word) {
    k_finalize_pub<<<1, 1024, 0, s>>>(nb, red.partial.p, scal, *pub);
    AN_CHECK_LAUNCH();
  } else {
    finalize(red, nb, 1, scal + S_RR, s);
  }
}

// z = (double) z32 with the Krylov products in FP64: r.z (and z.q)
template <int DOT>
__global__ __launch_bounds__(kBlock) void k_z_from32(int n, const float* __restrict__ z32,
                                                     const double* __restrict__ r,
                                                     const double* __restrict__ q,
                                                     double* __restrict__ z, RedArgs ra) {
  __shared__ double sh[2 * kWaves];
  double d[2] = {0.0, 0.0};
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    const double zi = (double)z32[i];
    z[i] = zi;
    d[0] += r[i] * zi;
    if (DOT == 2) d[1] += zi * q[i];
  }
  block_sums<DOT>(d, sh);
  store_partials<DOT>(d, blockIdx.x, ra);
}

void z_from_fp32(int n, const float* z32, const double* r, const double* q, double* z,
                 const RedArgs& ra, hipStream_t s) {
  const int nb = red_blocks(n);
  if (q)
    k_z_from32<2><<<nb, kBlock, 0, s>>>(n, z32, r, q, z, ra);
  else
    k_z_from32<1><<<nb, kBlock, 0, s>>>(n, z32, r, q, z, ra);
  AN_CHECK_LAUNCH();
  finalize_parts(ra, nb, q ? 2 : 1, s);
}

template <typename T>
void pcg_update(int n, T* x, T* r, const T* p, const T* q, double* scal, int num, int den,
                Reducer& red, hipStream_t s, const T* omd, T* y, const PcgPub* pub) {
  const int nb = red_blocks(n);
  k_pcg_update<T><<<nb, kBlock, 0, s>>>(n, x, r, p, q, scal, num, den, red.args(scal + S_RR), omd,
                                        y);
  AN_CHECK_LAUNCH();
  if (pub && pub->word) {
    k_finalize_pub<<<1, 1024, 0, s>>>(nb, red.partial.p, scal, *pub);
    AN_CHECK_LAUNCH();
  } else {
    finalize(red, nb, 1, scal + S_RR, s);
  }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_pcg_dir(int n, T* __restrict__ p,
                                                    const T* __restrict__ z,
                                                    const double* __restrict__ scal, int num,
                                                    int den, int first, double sign) {
  const double beta = first ? 0.0 : sign * scal[num] / scal[den];
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
    p[i] = first ? z[i] : (T)((double)z[i] + beta * (double)p[i]);
}

template <typename T>
void pcg_direction(int n, T* p, const T* z, double* scal, int num, int den, bool first,
                   hipStream_t s, double sign) {
  k_pcg_dir<T><<<grid_for(n, kBlock), kBlock, 0, s>>>(n, p, z, scal, num, den, first ? 1 : 0,
                                                      sign);
  AN_CHECK_LAUNCH();
}

__global__ void k_d2f(int n, const double* __restrict__ a, float* __restrict__ b) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}
__global__ void k_axpy_f2d(int n, double* __restrict__ x, const float* __restrict__ e) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] += (double)e[i];
}
__global__ void k_fill_zero(unsigned int* __restrict__ p, size_t nwords) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nwords) p[i] = 0u;
}
void copy_d2f(int n, const double* a, float* b, hipStream_t s) {
  k_d2f<<<grid_for(n, kBlock), kBlock, 0, s>>>(n, a, b);
  AN_CHECK_LAUNCH();
}
void axpy_f2d(int n, double* x, const float* e, hipStream_t s) {
  k_axpy_f2d<<<grid_for(n, kBlock), kBlock, 0, s>>>(n, x, e);
  AN_CHECK_LAUNCH();
}
void fill_zero(void* p, size_t bytes, hipStream_t s) {
  size_t words = bytes / 4;
  if (words == 0) return;
  k_fill_zero<<<grid_for((long long)words, kBlock), kBlock, 0, s>>>((unsigned int*)p, words);
  AN_CHECK_LAUNCH();
}

// explicit instantiations
template void spmv<double, double>(const DevCSR<double>&, const double*, double*, const double*,
                                   double*, int, Reducer&, hipStream_t, hipEvent_t, hipEvent_t);
template void spmv<float, float>(const DevCSR<float>&, const float*, float*, const float*, double*,
                                 int, Reducer&, hipStream_t, hipEvent_t, hipEvent_t);
template int residual_bb<double, double>(const DevCSR<double>&, const double*, const double*,
                                         double*, Reducer&, hipStream_t);
template int spmv_self_dot<double, double>(const DevCSR<double>&, const double*, double*,
                                           Reducer&, hipStream_t);
template void residual<double, double>(const DevCSR<double>&, const double*, const double*,
                                       double*, double*, Reducer&, hipStream_t);
template void residual<double, float>(const DevCSR<double>&, const double*, const double*, float*,
                                      double*, Reducer&, hipStream_t);
template void dot<double>(int, const double*, const double*, double*, int, Reducer&, hipStream_t);
template void dot<float>(int, const float*, const float*, double*, int, Reducer&, hipStream_t);
template void dot2<double>(int, const double*, const double*, const double*, double*, int,
                           Reducer&, hipStream_t);
template void dot2<float>(int, const float*, const float*, const float*, double*, int, Reducer&,
                          hipStream_t);
template void pcg_update<double>(int, double*, double*, const double*, const double*, double*, int,
                                 int, Reducer&, hipStream_t, const double*, double*,
                                 const PcgPub*);
template void pcg_update<float>(int, float*, float*, const float*, const float*, double*, int, int,
                                Reducer&, hipStream_t, const float*, float*, const PcgPub*);
template void pcg_direction<double>(int, double*, const double*, double*, int, int, bool,
                                    hipStream_t, double);
template void pcg_direction<float>(int, float*, const float*, double*, int, int, bool,
                                   hipStream_t, double);

}  // namespace an
