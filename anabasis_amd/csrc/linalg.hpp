// linalg.hpp -- sparse/vector primitives with device-resident scalars.
#pragma once

#include "an_internal.hpp"

namespace an {

// Row blocks of the CSR-stream kernels: consecutive rows grouped so that a
// block holds at most kStreamRows rows and at most kStreamCap nonzeros (a
// single longer row forms a block of its own).  One 64-lane wave owns one
// block.  start[b] = {first row, its first nonzero}; start[count] = {n, nnz}.
// 512 keeps 64 rows of up to 8 nonzeros in one block while the staging
// registers (8 nonzeros per lane) and LDS (4 KB per wave) allow 8 waves per
// SIMD: on the 2.1 M-row pressure matrix the SpMV ran at 48 us with 768
// (5 waves per SIMD) and 39 us with 512.
static constexpr int kStreamRows = 64;
static constexpr int kStreamCap = 512;
static constexpr int kStreamPer = kStreamCap / 64;  // staged nonzeros per lane
struct RowBlocks {
  DevBuf<int2> start;
  int count = 0;
};
// Greedy partition of rowOff[0..n] (synchronises once to read the count).
void build_row_blocks(int n, const int* rowOff, RowBlocks& rb, hipStream_t s);

// Column offsets of the CSR-stream kernels: c16[k] = cols[k] - r0, r0 the
// first row of the row block holding nonzero k, as 16-bit integers, when
// every offset of the matrix fits (|col - row| below about 32 700: true of
// banded mesh orderings); returns false, leaving c16 empty, otherwise.  The
// streaming kernels then read 2 bytes of column index per nonzero instead
// of 4.  Synchronises once.
bool build_col_offsets(long long nnz, const int* cols, const RowBlocks& rb, DevBuf<int16_t>& c16,
                       hipStream_t s);

// Device CSR matrix (values in T).  c16 (with has16) holds the column
// offsets above; the int32 cols stay for every other use.
template <typename T>
struct DevCSR {
  int n = 0;
  long long nnz = 0;
  DevBuf<int> rowOff, cols, diag;
  DevBuf<T> vals;
  RowBlocks rb;
  DevBuf<int16_t> c16;
  bool has16 = false;
  // offsets16 false: no column offsets (callers that size their memory by
  // rows only)
  void finish_pattern(hipStream_t s, bool offsets16 = true) {
    build_row_blocks(n, rowOff.p, rb, s);
    has16 = offsets16 && build_col_offsets(nnz, cols.p, rb, c16, s);
  }
};

// Column of nonzero k from the 16-bit offsets of its row block (first row
// r0); indexes like the int32 column array it stands for.
struct ColOffsets {
  const int16_t* __restrict__ p;
  int r0;
  __device__ __forceinline__ int operator[](int k) const { return r0 + (int)p[k]; }
};

// Waves synchronise their own LDS slice without a block barrier.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Row block of wave w of (remapped) block lb: rows [r0, r1), nonzeros
// [k0, k1); stream = the block's products fit the LDS slice.
struct WaveRows {
  int r0, r1, k0, k1;
  bool stream;
};
__device__ __forceinline__ WaveRows wave_rows(int nrb, const int2* __restrict__ rbs, int gw) {
  WaveRows q{0, 0, 0, 0, false};
  if (gw < nrb) {
    const int2 a = rbs[gw], b = rbs[gw + 1];
    q.r0 = a.x;
    q.r1 = b.x;
    q.k0 = a.y;
    q.k1 = b.y;
    q.stream = (q.k1 - q.k0) <= kStreamCap;
  }
  return q;
}

// Stage prod[k - k0] = vals[k] * xval(cols[k]) for the block's nonzeros:
// every lane first issues all of its (at most kStreamPer) index/value
// loads, then all of its gathers, so a wave waits on two memory round trips
// rather than one per nonzero.  Slots past the block's end re-read its last
// nonzero (no divergent branches between the loads, so the compiler keeps
// them all in flight) and are not stored.  Requires k1 > k0.
// cols: the int32 column array or the block's ColOffsets.
template <typename VT, typename CI, typename XF>
__device__ __forceinline__ void stage_products(const WaveRows& q, int lane, CI cols,
                                               const VT* __restrict__ vals, XF xval,
                                               double* __restrict__ prod) {
  int c[kStreamPer];
  double v[kStreamPer], xv[kStreamPer];
#pragma unroll
  for (int u = 0; u < kStreamPer; ++u) {
    const int k = min(q.k0 + lane + 64 * u, q.k1 - 1);
    c[u] = cols[k];
    v[u] = (double)vals[k];
  }
#pragma unroll
  for (int u = 0; u < kStreamPer; ++u) xv[u] = xval(c[u]);
#pragma unroll
  for (int u = 0; u < kStreamPer; ++u) {
    const int k = q.k0 + lane + 64 * u;
    if (k < q.k1) prod[k - q.k0] = v[u] * xv[u];
  }
}

// NC block sums, each in the tree order of a single block sum (valid in
// thread 0 afterwards); sh holds NC * (blockDim.x / 64) doubles.
template <int NC>
__device__ __forceinline__ void block_sums(double* v, double* sh) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = (int)(blockDim.x >> 6);
#pragma unroll
  for (int c = 0; c < NC; ++c)
    for (int off = 32; off > 0; off >>= 1) v[c] += __shfl_down(v[c], off, 64);
  __syncthreads();
  if (lane == 0)
#pragma unroll
    for (int c = 0; c < NC; ++c) sh[c * nw + w] = v[c];
  __syncthreads();
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    double t = 0.0;
    if (threadIdx.x == 0)
      for (int i = 0; i < nw; ++i) t += sh[c * nw + i];
    v[c] = t;
  }
}

// Scalar slots of a solver (device memory).
enum ScalarSlot { S_PQ = 0, S_RR = 1, S_RHO0 = 2, S_RHO1 = 3, S_BB = 4, S_TMP = 5, S_TMP2 = 6,
                  S_AL = 7, S_OM = 8, S_R0V = 9, S_TS = 10, S_TT = 11, S_SS = 12,
                  S_FRZ = 13, S_FZQ = 14,  // flexible CG: r.z, z.q
                  S_PROJ = 16,             // 16 projection coefficients
                  S_NU = 32,               // ||A w||^2 of a new direction, then ||A x||^2
                  S_RB = 34,               // a solve's first r.r and b.b
                  S_COUNT = 40 };

// Deterministic reduction workspace: per-block partials, summed in index
// order by a one-block finalize kernel.
// A reducing kernel's view of it: where its block partials go and where
// the finalize kernel that follows it puts the sums.
struct RedArgs {
  double* partial;
  double* out;
};
struct Reducer {
  DevBuf<double> partial;  // 4 doubles per block
  int cap = 0;
  void init(int maxBlocks, hipStream_t s);
  RedArgs args(double* out) const { return RedArgs{partial.p, out}; }
};

// Block partials of NC sums (bs, valid in thread 0) -> partial[blockId*NC
// + c], for finalize_parts() to add in index order.  (A last-block finish
// inside the reducing kernel would need a device-scope fence and one
// contended ticket per block: far slower here than a separate launch.)
template <int NC>
__device__ __forceinline__ void store_partials(const double* bs, int blockId, const RedArgs& ra) {
  if (threadIdx.x == 0)
#pragma unroll
    for (int c = 0; c < NC; ++c) ra.partial[(size_t)blockId * NC + c] = bs[c];
}

// out[c] = sum_i partial[i*comps + c], i < nparts (one block, fixed order)
void finalize(const Reducer& red, int nparts, int comps, double* out, hipStream_t s);
void finalize_parts(const RedArgs& ra, int nparts, int comps, hipStream_t s);

// y = A x.  When dotWith != nullptr also computes dot(dotWith, y) into
// scal[slot] (deterministically).  evStart/evStop, when given, record the
// start and end of the SpMV kernel itself (hipExtLaunchKernel), excluding
// the reduction's finalize kernel.
template <typename VT, typename XT>
void spmv(const DevCSR<VT>& A, const XT* x, XT* y, const XT* dotWith, double* scal, int slot,
          Reducer& red, hipStream_t s, hipEvent_t evStart = nullptr, hipEvent_t evStop = nullptr);
// scal[slot] = dot(a, b)
template <typename T>
void dot(int n, const T* a, const T* b, double* scal, int slot, Reducer& red, hipStream_t s);

// scal[slot] = a.b, scal[slot + 1] = b.c
template <typename T>
void dot2(int n, const T* a, const T* b, const T* c, double* scal, int slot, Reducer& red,
          hipStream_t s);

// Outcome of one PCG iteration for the host, decided on the device by the
// kernel that finishes r.r: CONTINUE, CONVERGED (||r|| <= tol) or BREAKDOWN
// (p.q <= 0 or ||r|| not finite), checked in that order as HYPRE's PCG does.
enum PcgDecision { PCG_CONTINUE = 0, PCG_CONVERGED = 1, PCG_BREAKDOWN = 2 };
struct PcgPub {
  unsigned long long* word = nullptr;  // device view of coherent host memory
  unsigned stamp = 0;                  // iteration stamp (low 30 bits used)
  double tol = 0.0;
  double* hist = nullptr;              // device: hist[idx] = ||r|| (FP64)
  int idx = 0;
};
// The published word: stamp in bits 63..34, decision in 33..32 and the
// float value of ||r|| in 31..0 (a hint for the host's speculation only;
// stop decisions and the residual history use the FP64 value).  One
// naturally aligned 64-bit store, so the host never sees a stamp with
// another iteration's decision.
__host__ __device__ inline unsigned long long pcg_word(unsigned stamp, int code, float rn) {
  unsigned bits;
  __builtin_memcpy(&bits, &rn, sizeof bits);
  return ((unsigned long long)(stamp & 0x3fffffffu) << 34) |
         ((unsigned long long)(code & 3) << 32) | bits;
}

// x += alpha p ; r -= alpha q ; scal[S_RR] = r.r, alpha = scal[num]/scal[den]
// (alpha = 0 when scal[den] <= 0 or the quotient is not finite: a breakdown
// leaves x at the last good iterate, as HYPRE stops before the update);
// with y: also y = omd o r (the preconditioner's first Jacobi sweep).
// With pub.word: the kernel that finishes r.r also decides the iteration's
// outcome, stores ||r|| into pub.hist[pub.idx] and publishes
// pcg_word(stamp, decision, ||r||) with a system-scope release store; the
// host waits for the stamp instead of a copy and a synchronisation.
template <typename T>
void pcg_update(int n, T* x, T* r, const T* p, const T* q, double* scal, int num, int den,
                Reducer& red, hipStream_t s, const T* omd = nullptr, T* y = nullptr,
                const PcgPub* pub = nullptr);
// pcg_update<double> for an FP32 preconditioner: also r32 = (float) r and
// y32 = omd32 o r32 (the FP32 cycle's input and first sweep)
void pcg_update_p2(int n, double* x, double* r, const double* p, const double* q, double* scal,
                   int num, int den, Reducer& red, hipStream_t s, const float* omd32, float* r32,
                   float* y32, const PcgPub* pub);
// z = (double) z32; ra.out[0] = r.z and, with q, ra.out[1] = z.q (FP64)
void z_from_fp32(int n, const float* z32, const double* r, const double* q, double* z,
                 const RedArgs& ra, hipStream_t s);
// p = z + beta p, beta = sign * scal[num]/scal[den]  (first: p = z)
template <typename T>
void pcg_direction(int n, T* p, const T* z, double* scal, int num, int den, bool first,
                   hipStream_t s, double sign = 1.0);
// r = b - A x with the block partials of r.r and b.b (two per block) left
// in red.partial; returns the number of blocks (the caller finalizes)
template <typename VT, typename RT>
int residual_bb(const DevCSR<VT>& A, const double* b, const double* x, RT* r, Reducer& red,
                hipStream_t s);
// y = A x with the block partials of y.y left in red.partial; returns the
// number of blocks
template <typename VT, typename XT>
int spmv_self_dot(const DevCSR<VT>& A, const XT* x, XT* y, Reducer& red, hipStream_t s);
// y = A x with the block partials of dw.y and y.y (two per block) left in
// red.partial; returns the number of blocks
int spmv_dot_self(const DevCSR<double>& A, const double* x, double* y, const double* dw,
                  Reducer& red, hipStream_t s);
// r = b - A x ; scal[S_RR] = r.r (FP64 residual, output converted to RT)
template <typename VT, typename RT>
void residual(const DevCSR<VT>& A, const double* b, const double* x, RT* r, double* scal,
              Reducer& red, hipStream_t s);

void copy_d2f(int n, const double* a, float* b, hipStream_t s);
void axpy_f2d(int n, double* x, const float* e, hipStream_t s);  // x += e
void fill_zero(void* p, size_t bytes, hipStream_t s);  // kernel fill (graph-safe)

}  // namespace an
