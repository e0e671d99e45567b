// linalg.hip -- CSR SpMV, fused Krylov vector updates and deterministic
// reductions for the PCG/BiCGSTAB solvers.
//
// SpMV (CSR-stream, wave granularity): each 64-lane wavefront owns 64
// consecutive rows.  The lanes first stream the rows' nonzeros with fully
// coalesced loads of vals/cols (lane k reads nonzero k0+k), multiply by the
// gathered x and stage the products in LDS; then lane i sums row i's
// products from LDS in CSR order.  Rows are 4-27 entries long on polyhedral
// meshes, so thread-per-row loads would be strided by the row length; the
// stream form turns them into unit-stride loads.  A wave whose rows hold more
// nonzeros than its LDS slice falls back to thread-per-row.  Sums run in CSR
// order, so y equals the serial product bit for bit.
//
// Reductions: every block writes a partial sum (fixed element->block map,
// fixed shuffle tree), the last block to finish (atomic ticket) adds the
// partials in index order and writes the scalar to device memory: results
// are deterministic and scalars never leave the GPU inside an iteration.
#include "linalg.hpp"

namespace an {

static constexpr int kBlock = 256;
static constexpr int kWaves = kBlock / 64;
static constexpr int kCap = 768;  // staged products per wave (6 KB)
static constexpr int kRedBlocks = 1024;

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

// Writes this block's partial; the last block reduces all partials into *out.
__device__ void finish_reduction(double blockVal, int blockId, int nblocks,
                                 double* __restrict__ partial, unsigned int* __restrict__ counter,
                                 double* __restrict__ out, double* sh) {
  __shared__ int isLast;
  if (threadIdx.x == 0) {
    partial[blockId] = blockVal;
    __threadfence();
    unsigned int t = atomicAdd(counter, 1u);
    isLast = (t == (unsigned int)nblocks - 1u);
  }
  __syncthreads();
  if (!isLast) return;
  __threadfence();
  double s = 0.0;
  const volatile double* vp = partial;
  for (int i = threadIdx.x; i < nblocks; i += blockDim.x) s += vp[i];
  s = block_sum(s, sh);
  if (threadIdx.x == 0) {
    *out = s;
    *counter = 0u;
  }
}

void Reducer::init(int maxBlocks, hipStream_t s) {
  cap = maxBlocks;
  partial.alloc((size_t)maxBlocks);
  counter.alloc(1);
  counter.zero(s);
}

// ------------------------------------------------------------------ SpMV
// MODE 0: y = A x ; MODE 1: y = b - A x.  DOT 0: none; 1: dot(dw, y);
// 2: dot(y, y).
template <typename VT, typename XT, typename YT, int MODE, int DOT>
__global__ __launch_bounds__(kBlock) void k_spmv(int n, const int* __restrict__ rowOff,
                                                 const int* __restrict__ cols,
                                                 const VT* __restrict__ vals,
                                                 const XT* __restrict__ x,
                                                 const double* __restrict__ b, YT* __restrict__ y,
                                                 const YT* __restrict__ dw,
                                                 double* __restrict__ partial,
                                                 unsigned int* __restrict__ counter,
                                                 double* __restrict__ out) {
  __shared__ double prod[kWaves][kCap];
  __shared__ double sh[kWaves];
  const int lb = xcd_block(gridDim.x);
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int r0 = (lb * kWaves + w) * 64;
  const int r1 = min(r0 + 64, n);
  int k0 = 0, k1 = 0;
  bool stream = false;
  if (r0 < n) {
    k0 = rowOff[r0];
    k1 = rowOff[r1];
    stream = (k1 - k0) <= kCap;
  }
  if (stream)
    for (int k = k0 + lane; k < k1; k += 64) prod[w][k - k0] = (double)vals[k] * (double)x[cols[k]];
  __syncthreads();
  double dv = 0.0;
  const int row = r0 + lane;
  if (row < r1) {
    const int kb = rowOff[row], ke = rowOff[row + 1];
    double s = 0.0;
    if (stream)
      for (int k = kb; k < ke; ++k) s += prod[w][k - k0];
    else
      for (int k = kb; k < ke; ++k) s += (double)vals[k] * (double)x[cols[k]];
    YT v;
    if (MODE == 1)
      v = (YT)(b[row] - s);
    else
      v = (YT)s;
    y[row] = v;
    if (DOT == 1) dv = (double)dw[row] * (double)v;
    if (DOT == 2) dv = (double)v * (double)v;
  }
  if (DOT) {
    double bs = block_sum(dv, sh);
    finish_reduction(bs, lb, gridDim.x, partial, counter, out, sh);
  }
}

template <typename VT, typename XT>
void spmv(const DevCSR<VT>& A, const XT* x, XT* y, const XT* dotWith, double* scal, int slot,
          Reducer& red, hipStream_t s) {
  const int rowsPerBlock = kBlock;
  const int nb = grid_for(A.n, rowsPerBlock);
  if (dotWith) {
    if (nb > red.cap) throw Error(7, "reducer too small");
    k_spmv<VT, XT, XT, 0, 1><<<nb, kBlock, 0, s>>>(A.n, A.rowOff.p, A.cols.p, A.vals.p, x, nullptr,
                                                   y, dotWith, red.partial.p, red.counter.p,
                                                   scal + slot);
  } else {
    k_spmv<VT, XT, XT, 0, 0><<<nb, kBlock, 0, s>>>(A.n, A.rowOff.p, A.cols.p, A.vals.p, x, nullptr,
                                                   y, nullptr, nullptr, nullptr, nullptr);
  }
  AN_CHECK_LAUNCH();
}

template <typename VT, typename RT>
void residual(const DevCSR<VT>& A, const double* b, const double* x, RT* r, double* scal,
              Reducer& red, hipStream_t s) {
  const int nb = grid_for(A.n, kBlock);
  if (nb > red.cap) throw Error(7, "reducer too small");
  k_spmv<VT, double, RT, 1, 2><<<nb, kBlock, 0, s>>>(A.n, A.rowOff.p, A.cols.p, A.vals.p, x, b, r,
                                                     nullptr, red.partial.p, red.counter.p,
                                                     scal + S_RR);
  AN_CHECK_LAUNCH();
}

// ---------------------------------------------------------- vector ops
static inline int red_blocks(int n) {
  int nb = grid_for(n, kBlock);
  return nb < kRedBlocks ? nb : kRedBlocks;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_dot(int n, const T* __restrict__ a,
                                                const T* __restrict__ b,
                                                double* __restrict__ partial,
                                                unsigned int* __restrict__ counter,
                                                double* __restrict__ out) {
  __shared__ double sh[kWaves];
  double v = 0.0;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
    v += (double)a[i] * (double)b[i];
  double bs = block_sum(v, sh);
  finish_reduction(bs, blockIdx.x, gridDim.x, partial, counter, out, sh);
}

template <typename T>
void dot(int n, const T* a, const T* b, double* scal, int slot, Reducer& red, hipStream_t s) {
  k_dot<T><<<red_blocks(n), kBlock, 0, s>>>(n, a, b, red.partial.p, red.counter.p, scal + slot);
  AN_CHECK_LAUNCH();
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_pcg_update(int n, T* __restrict__ x,
                                                       T* __restrict__ r,
                                                       const T* __restrict__ p,
                                                       const T* __restrict__ q,
                                                       const double* __restrict__ scal, int num,
                                                       int den, double* __restrict__ partial,
                                                       unsigned int* __restrict__ counter,
                                                       double* __restrict__ out) {
  __shared__ double sh[kWaves];
  const double alpha = scal[num] / scal[den];
  double v = 0.0;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    x[i] = (T)((double)x[i] + alpha * (double)p[i]);
    const T ri = (T)((double)r[i] - alpha * (double)q[i]);
    r[i] = ri;
    v += (double)ri * (double)ri;
  }
  double bs = block_sum(v, sh);
  finish_reduction(bs, blockIdx.x, gridDim.x, partial, counter, out, sh);
}

template <typename T>
void pcg_update(int n, T* x, T* r, const T* p, const T* q, double* scal, int num, int den,
                Reducer& red, hipStream_t s) {
  k_pcg_update<T><<<red_blocks(n), kBlock, 0, s>>>(n, x, r, p, q, scal, num, den, red.partial.p,
                                                   red.counter.p, scal + S_RR);
  AN_CHECK_LAUNCH();
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_pcg_dir(int n, T* __restrict__ p,
                                                    const T* __restrict__ z,
                                                    const double* __restrict__ scal, int num,
                                                    int den, int first) {
  const double beta = first ? 0.0 : scal[num] / scal[den];
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
    p[i] = first ? z[i] : (T)((double)z[i] + beta * (double)p[i]);
}

template <typename T>
void pcg_direction(int n, T* p, const T* z, double* scal, int num, int den, bool first,
                   hipStream_t s) {
  k_pcg_dir<T><<<grid_for(n, kBlock), kBlock, 0, s>>>(n, p, z, scal, num, den, first ? 1 : 0);
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
                                   double*, int, Reducer&, hipStream_t);
template void spmv<float, float>(const DevCSR<float>&, const float*, float*, const float*, double*,
                                 int, Reducer&, hipStream_t);
template void residual<double, double>(const DevCSR<double>&, const double*, const double*,
                                       double*, double*, Reducer&, hipStream_t);
template void residual<double, float>(const DevCSR<double>&, const double*, const double*, float*,
                                      double*, Reducer&, hipStream_t);
template void dot<double>(int, const double*, const double*, double*, int, Reducer&, hipStream_t);
template void dot<float>(int, const float*, const float*, double*, int, Reducer&, hipStream_t);
template void pcg_update<double>(int, double*, double*, const double*, const double*, double*, int,
                                 int, Reducer&, hipStream_t);
template void pcg_update<float>(int, float*, float*, const float*, const float*, double*, int, int,
                                Reducer&, hipStream_t);
template void pcg_direction<double>(int, double*, const double*, double*, int, int, bool,
                                    hipStream_t);
template void pcg_direction<float>(int, float*, const float*, double*, int, int, bool,
                                   hipStream_t);

}  // namespace an
