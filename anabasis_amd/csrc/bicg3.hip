// bicg3.hip -- BiCGSTAB + Jacobi for three right-hand sides sharing one
// matrix (the u/v/w momentum systems of generic_simple_v1, which reuses the
// x-momentum matrix for y and z: main.cu:3401-3405).
//
// Vectors are interleaved, element (i, c) at 3i + c, so one SpMV streams the
// matrix once and gathers 24 contiguous bytes per nonzero for all three
// columns.  Every column keeps its own scalars (slot-major: sc[3*slot + c])
// and its own stop test; a converged column is frozen by the active mask.
//
// Per column the arithmetic is the single-RHS bicgstab_device's, operation
// for operation: same kernels' element->block maps (grid-stride loops over
// rows, CSR row blocks), the same shuffle trees and the same index-order
// finalize.  A column's iterates, scalars and iteration count are therefore
// bit-identical to a separate solve of that column.
#include <hip/hip_ext.h>

#include <algorithm>
#include <chrono>
#include <cmath>

#include "solver.hpp"

namespace an {

static constexpr int kB = 256;
static constexpr int kW = kB / 64;
static constexpr int kCap = kStreamCap;

__device__ __forceinline__ double wsum3(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// three block sums with the single-RHS tree (valid in thread 0)
__device__ __forceinline__ void block_sum3(double v[3], double (*sh)[kW]) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int c = 0; c < 3; ++c) v[c] = wsum3(v[c]);
  __syncthreads();
  if (lane == 0)
#pragma unroll
    for (int c = 0; c < 3; ++c) sh[c][w] = v[c];
  __syncthreads();
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    double t = 0.0;
    if (threadIdx.x == 0)
      for (int i = 0; i < (int)(blockDim.x >> 6); ++i) t += sh[c][i];
    v[c] = t;
  }
}

__device__ __forceinline__ void store3(const double v[3], int blk, double* __restrict__ partial) {
  if (threadIdx.x == 0)
#pragma unroll
    for (int c = 0; c < 3; ++c) partial[3 * blk + c] = v[c];
}

__global__ void k3_pack(int n, const double* __restrict__ a0, const double* __restrict__ a1,
                        const double* __restrict__ a2, double* __restrict__ X) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  X[3 * i] = a0[i];
  X[3 * i + 1] = a1[i];
  X[3 * i + 2] = a2[i];
}

__global__ void k3_unpack(int n, const double* __restrict__ X, double* __restrict__ a0,
                          double* __restrict__ a1, double* __restrict__ a2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  a0[i] = X[3 * i];
  a1[i] = X[3 * i + 1];
  a2[i] = X[3 * i + 2];
}

// The live column mask of a solve; without a state (the set-up kernels
// before the first stop test) all three columns.
__device__ __forceinline__ int live_mask(const Bicg3State* st) { return st ? st->mask : 7; }

// partial[3 blk + c] = block's share of sum_i A[3i+c] B[3i+c] (k_dot's map);
// nothing once no column is live
// COPY: B is not read but written, B = A, and the sums are A.A (the bits of
// a separate copy and dot of the two).
template <bool COPY = false>
__global__ __launch_bounds__(kB) void k3_dot(int n, const double* __restrict__ A,
                                             double* __restrict__ B,
                                             const Bicg3State* st,
                                             double* __restrict__ partial) {
  __shared__ double sh[3][kW];
  if (!live_mask(st)) return;  // grid-uniform
  double v[3] = {0.0, 0.0, 0.0};
  const int S = gridDim.x * kB;
  for (int i = blockIdx.x * kB + threadIdx.x; i < n; i += 2 * S) {
    const int j = (i + S < n) ? i + S : i;  // second element of the pair (or i again)
    double a[2][3], b[2][3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      a[0][c] = A[3 * i + c];
      a[1][c] = A[3 * j + c];
      if (COPY) {
        b[0][c] = a[0][c];
        b[1][c] = a[1][c];
        B[3 * i + c] = a[0][c];
        B[3 * j + c] = a[1][c];
      } else {
        b[0][c] = B[3 * i + c];
        b[1][c] = B[3 * j + c];
      }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) v[c] += a[0][c] * b[0][c];
    if (j != i)
#pragma unroll
      for (int c = 0; c < 3; ++c) v[c] += a[1][c] * b[1][c];
  }
  block_sum3(v, sh);
  store3(v, blockIdx.x, partial);
}

// MODE 0: Y = A X; MODE 1: Y = B - A X.  DOT 0: none; 1: dot(DW, Y);
// 2: dot(Y, Y); 3: dot(DW, Y) and dot(Y, Y) into partial[6 blk + c] and
// partial[6 blk + 3 + c] (per column, k_spmv's row blocks and trees).
// Columns outside the live mask are neither written nor reduced.
// C16: columns from the 16-bit offsets c16 (ColOffsets) instead of cols.
template <int MODE, int DOT, bool C16>
__global__ __launch_bounds__(kB) void k3_spmv(int nrb, const int2* __restrict__ rbs,
                                              const int* __restrict__ rowOff,
                                              const int* __restrict__ cols,
                                              const int16_t* __restrict__ c16,
                                              const double* __restrict__ vals,
                                              const double* __restrict__ X,
                                              const double* __restrict__ B,
                                              double* __restrict__ Y,
                                              const double* __restrict__ DW,
                                              const Bicg3State* st,
                                              double* __restrict__ partial) {
  __shared__ double prod[kW][kCap];
  __shared__ double sh[3][kW];
  const int mask = live_mask(st);
  if (!mask) return;  // grid-uniform
  const int lb = xcd_block(gridDim.x);
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const WaveRows q = wave_rows(nrb, rbs, lb * kW + w);
  const int row = q.r0 + lane;
  // the row's own operands are loaded with the row offsets, so their latency
  // hides behind the nonzero stream
  int kb = 0, ke = 0;
  double bw[3] = {0.0, 0.0, 0.0};
  if (row < q.r1) {
    kb = rowOff[row];
    ke = rowOff[row + 1];
    if (MODE == 1 || DOT == 1 || DOT == 3)
#pragma unroll
      for (int c = 0; c < 3; ++c) bw[c] = (MODE == 1) ? B[3 * row + c] : DW[3 * row + c];
  }
  static_assert(!(MODE == 1 && (DOT == 1 || DOT == 3)), "one row operand per launch");
  double dv[3] = {0.0, 0.0, 0.0}, dt[3] = {0.0, 0.0, 0.0};
  const bool stream = q.stream && q.k1 > q.k0;
  int cidx[kStreamPer];
  double v[kStreamPer], x3[kStreamPer][3];
  if (stream) {
#pragma unroll
    for (int u = 0; u < kStreamPer; ++u) {
      const int k = min(q.k0 + lane + 64 * u, q.k1 - 1);
      cidx[u] = C16 ? q.r0 + (int)c16[k] : cols[k];
      v[u] = vals[k];
    }
#pragma unroll
    for (int u = 0; u < kStreamPer; ++u)
#pragma unroll
      for (int c = 0; c < 3; ++c) x3[u][c] = X[3 * cidx[u] + c];
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    if (!((mask >> c) & 1)) continue;  // wave-uniform
    if (stream)
#pragma unroll
      for (int u = 0; u < kStreamPer; ++u) {
        const int k = q.k0 + lane + 64 * u;
        if (k < q.k1) prod[w][k - q.k0] = v[u] * x3[u][c];
      }
    wave_sync();
    if (row < q.r1) {
      double s = 0.0;
      if (stream)
        for (int k = kb; k < ke; ++k) s += prod[w][k - q.k0];
      else
        for (int k = kb; k < ke; ++k)
          s += vals[k] * X[3 * (C16 ? q.r0 + (int)c16[k] : cols[k]) + c];
      const double yv = (MODE == 1) ? bw[c] - s : s;
      Y[3 * row + c] = yv;
      if (DOT == 1 || DOT == 3) dv[c] = bw[c] * yv;
      if (DOT == 2) dv[c] = yv * yv;
      if (DOT == 3) dt[c] = yv * yv;
    }
    wave_sync();  // the slice is reused by the next column
  }
  if (DOT == 3) {
    block_sum3(dv, sh);
    block_sum3(dt, sh);
    if (threadIdx.x == 0)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        partial[6 * lb + c] = dv[c];
        partial[6 * lb + 3 + c] = dt[c];
      }
  } else if (DOT) {
    block_sum3(dv, sh);
    store3(dv, lb, partial);
  }
}

// k3_spmv with 16-bit column offsets when the matrix has them; with
// evStart/evStop the launch is timed (hipExtLaunchKernel)
template <int MODE, int DOT>
static void launch_spmv3(const DevCSR<double>& A, int nb, hipStream_t s, const double* X,
                         const double* B, double* Y, const double* DW, const Bicg3State* st,
                         double* partial, hipEvent_t evStart = nullptr,
                         hipEvent_t evStop = nullptr) {
  const int2* rbs = A.rb.start.p;
  const int* ro = A.rowOff.p;
  const int* ci = A.cols.p;
  const int16_t* c16 = A.c16.p;
  const double* va = A.vals.p;
  if (A.has16) {
    if (evStart)
      hipExtLaunchKernelGGL(k3_spmv<MODE, DOT, true>, dim3(nb), dim3(kB), 0, s, evStart, evStop,
                            0, A.rb.count, rbs, ro, ci, c16, va, X, B, Y, DW, st, partial);
    else
      k3_spmv<MODE, DOT, true><<<nb, kB, 0, s>>>(A.rb.count, rbs, ro, ci, c16, va, X, B, Y, DW,
                                                 st, partial);
  } else {
    if (evStart)
      hipExtLaunchKernelGGL(k3_spmv<MODE, DOT, false>, dim3(nb), dim3(kB), 0, s, evStart, evStop,
                            0, A.rb.count, rbs, ro, ci, c16, va, X, B, Y, DW, st, partial);
    else
      k3_spmv<MODE, DOT, false><<<nb, kB, 0, s>>>(A.rb.count, rbs, ro, ci, c16, va, X, B, Y, DW,
                                                  st, partial);
  }
}

// p = r + beta (p - omega v); ph = dinv p   (k_bicg_p per column)
__global__ __launch_bounds__(kB) void k3_bicg_p(int n, const double* __restrict__ R,
                                                double* __restrict__ P,
                                                const double* __restrict__ V,
                                                const double* __restrict__ dinv,
                                                double* __restrict__ PH,
                                                const double* __restrict__ sc, int rhoNew,
                                                int rhoOld, int first,
                                                const Bicg3State* __restrict__ st) {
  const int mask = st->mask;
  if (!mask) return;
  double beta[3], om[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    beta[c] = first ? 0.0
                    : (sc[3 * rhoNew + c] / sc[3 * rhoOld + c]) * (sc[3 * S_AL + c] / sc[3 * S_OM + c]);
    om[c] = first ? 0.0 : sc[3 * S_OM + c];
  }
  // every column's operands are loaded together (a frozen column's are
  // read and dropped): loads behind a per-column branch waited one by one
  for (int i = blockIdx.x * kB + threadIdx.x; i < n; i += gridDim.x * kB) {
    const double di = dinv[i];
    double rv[3], pv[3], vv[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      rv[c] = R[3 * i + c];
      pv[c] = first ? 0.0 : P[3 * i + c];
      vv[c] = first ? 0.0 : V[3 * i + c];
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      if (!((mask >> c) & 1)) continue;
      const int e = 3 * i + c;
      const double pi = first ? rv[c] : rv[c] + beta[c] * (pv[c] - om[c] * vv[c]);
      P[e] = pi;
      PH[e] = di * pi;
    }
  }
}

// alpha = rhoNew / r0v; s = r - alpha v; sh = dinv s; ss = s.s
__global__ __launch_bounds__(kB) void k3_bicg_s(int n, const double* __restrict__ R,
                                                const double* __restrict__ V,
                                                const double* __restrict__ dinv,
                                                double* __restrict__ S, double* __restrict__ SH,
                                                double* __restrict__ sc, int rhoNew,
                                                const Bicg3State* __restrict__ st,
                                                double* __restrict__ partial) {
  __shared__ double sh[3][kW];
  const int mask = st->mask;
  if (!mask) return;
  double alpha[3], acc[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int c = 0; c < 3; ++c) alpha[c] = sc[3 * rhoNew + c] / sc[3 * S_R0V + c];
  // a thread's grid-stride elements four at a time, all loads first, the
  // sums in element order (the bits of one element at a time)
  constexpr int U = 4;
  const int St = gridDim.x * kB;
  for (int i = blockIdx.x * kB + threadIdx.x; i < n; i += U * St) {
    int el[U];
    double di[U], r[U][3], v[U][3];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      el[u] = min(i + u * St, n - 1);
      di[u] = dinv[el[u]];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        r[u][c] = R[3 * el[u] + c];
        v[u][c] = V[3 * el[u] + c];
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (i + u * St >= n) break;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        if (!((mask >> c) & 1)) continue;
        const int e = 3 * el[u] + c;
        const double si = r[u][c] - alpha[c] * v[u][c];
        S[e] = si;
        SH[e] = di[u] * si;
        acc[c] += si * si;
      }
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0)
#pragma unroll
    for (int c = 0; c < 3; ++c)
      if ((mask >> c) & 1) sc[3 * S_AL + c] = alpha[c];
  block_sum3(acc, sh);
  store3(acc, blockIdx.x, partial);
}

// omega = ts/tt; x += alpha ph (+ omega sh); r = s (- omega t), with
// ph = dinv o p and sh = dinv o s formed here from p and s (the products
// k3_bicg_p and k3_bicg_s stored for the SpMV) (k_bicg_x per column).
// Partials partial[6 blk + c] = r.r and partial[6 blk + 3 + c] = the next
// rho = r0.r, accumulated in k3_dot's element order and tree over the same
// grid (the bits of a separate k3_dot); frozen columns add nothing.
__global__ __launch_bounds__(kB) void k3_bicg_x(int n, double* __restrict__ X,
                                                double* __restrict__ R,
                                                const double* __restrict__ P,
                                                const double* __restrict__ S,
                                                const double* __restrict__ T,
                                                const double* __restrict__ dinv,
                                                const double* __restrict__ R0,
                                                double* __restrict__ sc,
                                                const Bicg3State* __restrict__ st,
                                                double* __restrict__ partial) {
  __shared__ double sh[3][kW];
  const int mask = st->mask;
  if (!mask) return;
  double alpha[3], omega[3], acc[3] = {0.0, 0.0, 0.0}, rho[3] = {0.0, 0.0, 0.0};
  bool convS[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    alpha[c] = sc[3 * S_AL + c];
    const double tt = sc[3 * S_TT + c];
    omega[c] = (tt > 0.0) ? sc[3 * S_TS + c] / tt : 0.0;
    convS[c] = sqrt(sc[3 * S_SS + c]) <= st->tol[c];
  }
  // two grid-stride elements per step, loads first (three or four take
  // 132 or 169 VGPRs: fewer than the grid's four waves per SIMD fit)
  constexpr int U = 2;
  const int St = gridDim.x * kB;
  for (int i = blockIdx.x * kB + threadIdx.x; i < n; i += U * St) {
    int el[U];
    double di[U], xv[U][3], pv[U][3], sv[U][3], tv[U][3], r0[U][3];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      el[u] = min(i + u * St, n - 1);
      di[u] = dinv[el[u]];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const int e = 3 * el[u] + c;
        xv[u][c] = X[e];
        pv[u][c] = P[e];
        sv[u][c] = S[e];
        tv[u][c] = T[e];
        r0[u][c] = R0[e];
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (i + u * St >= n) break;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        if (!((mask >> c) & 1)) continue;
        const int e = 3 * el[u] + c;
        double ri;
        if (convS[c]) {
          X[e] = xv[u][c] + alpha[c] * (di[u] * pv[u][c]);
          ri = sv[u][c];
        } else {
          X[e] = xv[u][c] + (alpha[c] * (di[u] * pv[u][c]) + omega[c] * (di[u] * sv[u][c]));
          ri = sv[u][c] - omega[c] * tv[u][c];
        }
        R[e] = ri;
        acc[c] += ri * ri;
        rho[c] += r0[u][c] * ri;
      }
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0)
#pragma unroll
    for (int c = 0; c < 3; ++c)
      if ((mask >> c) & 1) sc[3 * S_OM + c] = omega[c];
  block_sum3(acc, sh);
  block_sum3(rho, sh);
  if (threadIdx.x == 0)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      partial[6 * blockIdx.x + c] = acc[c];
      partial[6 * blockIdx.x + 3 + c] = rho[c];
    }
}

__global__ void k3_dinv(int n, const int* __restrict__ ro, const int* __restrict__ ci,
                        const double* __restrict__ va, double* __restrict__ dinv) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double d = 0.0;
  for (int k = ro[i]; k < ro[i + 1]; ++k)
    if (ci[k] == i) d = va[k];
  dinv[i] = (d != 0.0) ? 1.0 / d : 1.0;
}

// k_finalize_nc<NC> (the same loads and tree, so the same bits) that does
// nothing once no column is live
template <int NC>
__global__ __launch_bounds__(1024) void k3_fin(int nparts, const double* __restrict__ partial,
                                               double* __restrict__ out,
                                               const Bicg3State* st) {
  __shared__ double sh[NC * 16];
  if (!live_mask(st)) return;
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

// The stop test before the first iteration: tolerances, b = 0 columns
// (x = 0, HYPRE's convention) and columns whose r0 already meets the tolerance.
__global__ void k3_init(const double* __restrict__ sc, Bicg3State* __restrict__ st, int activeIn,
                        double relTol, double absTol) {
  if (threadIdx.x != 0) return;
  int m = activeIn & 7, z = 0;
  for (int c = 0; c < 3; ++c) {
    const double bnorm = sqrt(sc[3 * S_BB + c]), rn = sqrt(sc[3 * S_RR + c]);
    const double tol = fmax(relTol * bnorm, absTol);
    st->bnorm[c] = bnorm;
    st->rn[c] = rn;
    st->tol[c] = tol;
    st->its[c] = 0;
    st->conv[c] = 0;
    if (!((m >> c) & 1)) continue;
    if (bnorm == 0.0) {
      z |= 1 << c;
      st->conv[c] = 1;
      m &= ~(1 << c);
    } else if (rn <= tol) {
      st->conv[c] = 1;
      m &= ~(1 << c);
    }
  }
  st->mask = m;
  st->zero = z;
}

__global__ void k3_zero_cols(int n, double* __restrict__ X, const Bicg3State* __restrict__ st) {
  const int z = st->zero, i = blockIdx.x * blockDim.x + threadIdx.x;
  if (!z || i >= n) return;
#pragma unroll
  for (int c = 0; c < 3; ++c)
    if ((z >> c) & 1) X[3 * i + c] = 0.0;
}

// k3_fin<6> for r.r and the next rho (into S_RR and rhoNext), then
// iteration 'it''s stop test per live column
// (converged: ||s|| or ||r|| <= tol; dropped: rho, r0.v or omega = 0 or ||r||
// not finite) and the publication of the outcome: pcg_word(stamp,
// CONTINUE while a column is live else CONVERGED, max ||r||/tol over the
// live columns as the host's speculation hint).
__global__ __launch_bounds__(1024) void k3_fin_decide(int nparts,
                                                      const double* __restrict__ partial,
                                                      double* __restrict__ sc,
                                                      Bicg3State* __restrict__ st, int it,
                                                      int rhoNew, int rhoNext,
                                                      unsigned long long* word,
                                                      unsigned stamp) {
  __shared__ double sh[6 * 16];
  const int mask = st->mask;
  if (!mask) {  // nothing was live: announce the stop again
    if (threadIdx.x == 0)
      __hip_atomic_store(word, pcg_word(stamp, PCG_CONVERGED, 0.0f), __ATOMIC_RELEASE,
                         __HIP_MEMORY_SCOPE_SYSTEM);
    return;
  }
  double v[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
  for (int i = threadIdx.x; i < nparts; i += blockDim.x)
#pragma unroll
    for (int c = 0; c < 6; ++c) v[c] += partial[(size_t)i * 6 + c];
  block_sums<6>(v, sh);
  if (threadIdx.x != 0) return;
  int m = mask;
  float worst = 0.0f;
  for (int c = 0; c < 3; ++c) {
    sc[3 * S_RR + c] = v[c];
    sc[3 * rhoNext + c] = v[3 + c];
    if (!((mask >> c) & 1)) continue;
    const double rn = sqrt(v[c]), tol = st->tol[c];
    st->rn[c] = rn;
    st->its[c] = it;
    if (sqrt(sc[3 * S_SS + c]) <= tol || rn <= tol) {
      st->conv[c] = 1;
      m &= ~(1 << c);
    } else if (sc[3 * rhoNew + c] == 0.0 || sc[3 * S_R0V + c] == 0.0 ||
               sc[3 * S_OM + c] == 0.0 || !isfinite(rn)) {
      m &= ~(1 << c);
    } else {
      worst = fmaxf(worst, tol > 0.0 ? (float)(rn / tol) : __builtin_huge_valf());
    }
  }
  st->mask = m;
  // release at system scope: the state written above is visible before
  // the word that announces it
  __hip_atomic_store(word, pcg_word(stamp, m ? PCG_CONTINUE : PCG_CONVERGED, worst),
                     __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

void Bicg3Work::ensure(int nn, int nbSpmv, hipStream_t st) {
  const int need = std::max(nbSpmv, 1024) * 2 + 16;  // 6 doubles per block fit in 4 * need
  if (n == nn && scal.p && red.cap >= need) return;
  n = nn;
  red.init(need, st);
  scal.alloc(3 * S_COUNT);
  state.alloc((sizeof(Bicg3State) + sizeof(double) - 1) / sizeof(double));
  dinv.alloc(n);
  for (DevBuf<double>* b : {&x, &b3, &r, &r0, &p, &v, &ph, &s_, &sh, &t}) b->alloc(3 * (size_t)n);
  if (!hs) AN_HIP(hipHostMalloc((void**)&hs, sizeof(Bicg3State)));
}

Bicg3Work::~Bicg3Work() {
  if (hs) (void)hipHostFree(hs);
}

// The host queues iteration k+1 before it waits for the published outcome
// of iteration k unless the last two hints predict that k ends the solve
// (the PCG rule with max ||r||/tol against 1); an iteration queued after
// the last live column stopped finds the mask empty and every kernel
// returns at once.
static inline bool speculate3(double q, double qPrev) {
  return !(qPrev > 0.0 && q * (q / qPrev) <= 1.0);
}

an_status bicgstab3_device(const DevCSR<double>& A, const double* const b[3], double* const x[3],
                           int maxIter, double relTol, double absTol, Bicg3Work& w,
                           hipStream_t s, an_solve_info info[3], int activeIn,
                           hipEvent_t evStart, hipEvent_t evStop, const double* const* x0,
                           const double* dinvIn) {
  const int n = A.n;
  auto t0 = std::chrono::steady_clock::now();
  const int nbS = grid_for(A.rb.count, kW);
  if (A.n > 0 && A.rb.count == 0) throw Error(7, "CSR row blocks not built");
  w.ensure(n, nbS, s);
  const int G = std::min(grid_for(n, kB), 1024);
  const int Gd = std::min(grid_for(n, kB), 1024);  // dot(): red_blocks
  const int gN = grid_for(n, kB);
  double* sc = w.scal.p;
  double* part = w.red.partial.p;
  Bicg3State* st = reinterpret_cast<Bicg3State*>(w.state.p);
  // the Jacobi scaling: the caller's (formed where the diagonal was
  // assembled) or 1 / the diagonal entry of each row
  if (!dinvIn) k3_dinv<<<gN, kB, 0, s>>>(n, A.rowOff.p, A.cols.p, A.vals.p, w.dinv.p);
  k3_pack<<<gN, kB, 0, s>>>(n, b[0], b[1], b[2], w.b3.p);
  // the start: x0 when given (x is then written only at the end)
  const double* const* xs = x0 ? x0 : x;
  k3_pack<<<gN, kB, 0, s>>>(n, xs[0], xs[1], xs[2], w.x.p);
  AN_CHECK_LAUNCH();
  // ||b||, r = b - A x, ||r||, the first stop test
  k3_dot<<<Gd, kB, 0, s>>>(n, w.b3.p, w.b3.p, nullptr, part);
  k3_fin<3><<<1, 1024, 0, s>>>(Gd, part, sc + 3 * S_BB, nullptr);
  launch_spmv3<1, 2>(A, nbS, s, w.x.p, w.b3.p, w.r.p, nullptr, nullptr, part);
  k3_fin<3><<<1, 1024, 0, s>>>(nbS, part, sc + 3 * S_RR, nullptr);
  k3_init<<<1, 64, 0, s>>>(sc, st, activeIn, relTol, absTol);
  k3_zero_cols<<<gN, kB, 0, s>>>(n, w.x.p, st);
  AN_CHECK_LAUNCH();
  // r0 = r and rho of the first iteration = r0.r in one pass; each later
  // rho comes out of k3_bicg_x
  k3_dot<true><<<Gd, kB, 0, s>>>(n, w.r.p, w.r0.p, st, part);
  k3_fin<3><<<1, 1024, 0, s>>>(Gd, part, sc + 3 * S_RHO1, st);
  AN_CHECK_LAUNCH();
  const unsigned base = w.stamp;
  const double* dinv = dinvIn ? dinvIn : w.dinv.p;
  auto queue_iteration = [&](int it) {
    const int rhoNew = (it & 1) ? S_RHO1 : S_RHO0;
    const int rhoOld = (it & 1) ? S_RHO0 : S_RHO1;
    k3_bicg_p<<<gN, kB, 0, s>>>(n, w.r.p, w.p.p, w.v.p, dinv, w.ph.p, sc, rhoNew, rhoOld,
                                it == 1 ? 1 : 0, st);
    launch_spmv3<0, 1>(A, nbS, s, w.ph.p, nullptr, w.v.p, w.r0.p, st, part,
                       it == 1 ? evStart : nullptr, evStop);
    k3_fin<3><<<1, 1024, 0, s>>>(nbS, part, sc + 3 * S_R0V, st);
    k3_bicg_s<<<G, kB, 0, s>>>(n, w.r.p, w.v.p, dinv, w.s_.p, w.sh.p, sc, rhoNew, st, part);
    k3_fin<3><<<1, 1024, 0, s>>>(G, part, sc + 3 * S_SS, st);
    launch_spmv3<0, 3>(A, nbS, s, w.sh.p, nullptr, w.t.p, w.s_.p, st, part);
    k3_fin<6><<<1, 1024, 0, s>>>(nbS, part, sc + 3 * S_TS, st);  // S_TS[0..2], S_TT[0..2]
    k3_bicg_x<<<G, kB, 0, s>>>(n, w.x.p, w.r.p, w.p.p, w.s_.p, w.t.p, dinv, w.r0.p, sc, st,
                               part);
    k3_fin_decide<<<1, 1024, 0, s>>>(G, part, sc, st, it, rhoNew, rhoOld, w.pub.d,
                                     base + (unsigned)it);
    AN_CHECK_LAUNCH();
  };
  if (maxIter >= 1) {
    queue_iteration(1);
    double q = 0.0, qPrev = 0.0;
    for (int it = 1;; ++it) {
      const bool spec = it < maxIter && speculate3(q, qPrev);
      if (spec) queue_iteration(it + 1);  // runs while the host waits
      const unsigned long long word =
          w.pub.wait(base + (unsigned)it, s, "BiCGSTAB iteration was not published");
      qPrev = q;
      q = PubWord::hint(word);
      if (PubWord::code(word) != PCG_CONTINUE || it == maxIter) break;
      if (!spec) queue_iteration(it + 1);
    }
    w.stamp = base + (unsigned)maxIter + 1;  // past every stamp this solve may publish
  }
  k3_unpack<<<gN, kB, 0, s>>>(n, w.x.p, x[0], x[1], x[2]);
  AN_CHECK_LAUNCH();
  AN_HIP(hipMemcpyAsync(w.hs, st, sizeof(Bicg3State), hipMemcpyDeviceToHost, s));
  AN_HIP(hipStreamSynchronize(s));
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  const Bicg3State& h = *w.hs;
  an_status status = AN_OK;
  an_solve_info inf[3] = {};
  for (int c = 0; c < 3; ++c) {
    if (!((activeIn >> c) & 1)) continue;
    const bool zero = (h.zero >> c) & 1;
    inf[c].iterations = h.its[c];
    inf[c].converged = h.conv[c];
    inf[c].finalResNorm = zero ? 0.0 : h.rn[c];
    inf[c].finalRelResNorm = zero ? 0.0 : h.rn[c] / h.bnorm[c];
    inf[c].solveSeconds = secs;
    inf[c].spmvIndexBytes = A.has16 ? 2 : 4;
    if (!h.conv[c]) status = AN_ENOCONV;
  }
  if (info)
    for (int c = 0; c < 3; ++c) info[c] = inf[c];
  return status;
}

}  // namespace an
