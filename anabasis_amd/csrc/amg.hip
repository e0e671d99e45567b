// amg.hip -- unsmoothed-aggregation AMG built and applied on the GPU.
//
// Replaces HYPRE BoomerAMG (LP/src/hypre_backend.cu:66-92, GSV1
// main.cu:2239-2260).  Algorithm (identical to the CPU oracle, so hierarchies
// agree bit for bit):
//   coupling s_ij = -a_ij; strong if s_ij > 0 and s_ij >= theta*max_k s_ik
//   6 synchronous handshake rounds: every unmatched row proposes its strongest
//   unmatched strong neighbour (ties: symmetric edge hash); mutual proposals
//   pair; leftovers join the pair of their strongest matched strong neighbour
//   or stay single; aggregates numbered by their smallest member (scan)
//   2 such passes per level (Galerkin aggregate sum in between)
//   coarse matrix = P^T A P summed per coarse row in member, then CSR order
//   V-cycle: one damped-Jacobi sweep from zero, residual, restriction,
//   recursion, prolongation, one damped-Jacobi sweep; dense inverse on the
//   coarsest level (FP64 in both precisions).
// Setup kernels are one thread per row; every list built with atomics is
// sorted before use, so the hierarchy is deterministic.
#include "amg.hpp"

#include <algorithm>
#include <cmath>

namespace an {

static constexpr int kB = 256;
static constexpr int kW = kB / 64;
static constexpr int kCap = kStreamCap;
// Levels of at most kSpreadRows rows launch 8x the blocks they need and let
// only the blocks on XCD 0 work (hardware block b runs on XCD b % 8), so a
// cycle's small kernels read what the previous one wrote from one L2.
// Measured on the 2.08 M-cell pressure hierarchy: V-cycle 422 -> 417 us;
// spreading the 24 k-row level as well was slower (467 us: its 229 blocks
// then share one XCD's 32 CUs); with four matchings per level, taking the
// 5.3 k-row level in as well (a limit of 8192) made no difference.
static constexpr int kSpreadRows = 4096;
static inline int spread_of(int n) { return n <= kSpreadRows ? 8 : 1; }

// Logical block of a launch that needs 'need' blocks: a grid of exactly
// 8 x need blocks is spread (blocks off XCD 0 return -1, they have no work);
// otherwise each XCD takes a contiguous range of blocks (xcd_block), so
// rows that share cache lines meet in one L2
__device__ __forceinline__ int spread_block(int need) {
  if ((int)gridDim.x == 8 * need) return (blockIdx.x & 7) ? -1 : (int)(blockIdx.x >> 3);
  return xcd_block(gridDim.x);
}
static constexpr int kGalCap = 256;  // max entries of one coarse row

__host__ __device__ __forceinline__ uint32_t edge_hash(uint32_t a, uint32_t b) {
  uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
  uint32_t x = lo * 0x9E3779B1u ^ (hi + 0x7F4A7C15u);
  x ^= x >> 16;
  x *= 0x85EBCA6Bu;
  x ^= x >> 13;
  x *= 0xC2B2AE35u;
  x ^= x >> 16;
  return x;
}

// --------------------------------------------------------- matching ---
__global__ void k_smax(int n, const int* __restrict__ ro, const int* __restrict__ ci,
                       const double* __restrict__ va, double* __restrict__ smax) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double m = 0.0;
  for (int k = ro[i]; k < ro[i + 1]; ++k) {
    if (ci[k] == i) continue;
    double s = -va[k];
    if (s > m) m = s;
  }
  smax[i] = m;
}

__global__ void k_fill_int(int n, int* __restrict__ a, int v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

// mode 0: propose among unmatched neighbours (rows that are unmatched);
// mode 1: absorption target among matched neighbours (rows unmatched).
__global__ void k_cand(int n, const int* __restrict__ ro, const int* __restrict__ ci,
                       const double* __restrict__ va, const double* __restrict__ smax,
                       double theta, const int* __restrict__ match, int* __restrict__ cand,
                       int mode) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (match[i] >= 0) {
    cand[i] = -1;
    return;
  }
  double bs = 0.0;
  uint32_t bh = 0;
  int best = -1;
  const double thr = theta * smax[i];
  for (int k = ro[i]; k < ro[i + 1]; ++k) {
    int j = ci[k];
    if (j == i) continue;
    if (mode == 0 ? (match[j] >= 0) : (match[j] < 0)) continue;
    double s = -va[k];
    if (!(s > 0.0) || s < thr) continue;
    uint32_t h = edge_hash((uint32_t)i, (uint32_t)j);
    if (best < 0 || s > bs || (s == bs && h > bh)) {
      best = j;
      bs = s;
      bh = h;
    }
  }
  cand[i] = best;
}

__global__ void k_match(int n, const int* __restrict__ cand, int* __restrict__ match) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int j = cand[i];
  if (j >= 0 && cand[j] == i) match[i] = j;
}

__global__ void k_rep(int n, const int* __restrict__ match, const int* __restrict__ cand2,
                      int* __restrict__ rep) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  rep[i] = match[i] >= 0 ? (i < match[i] ? 1 : 0) : (cand2[i] < 0 ? 1 : 0);
}

__global__ void k_assign(int n, const int* __restrict__ match, const int* __restrict__ cand2,
                         const int* __restrict__ id, int* __restrict__ agg) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int r;
  if (match[i] >= 0)
    r = i < match[i] ? i : match[i];
  else if (cand2[i] >= 0) {
    int j = cand2[i];
    r = match[j] < j ? match[j] : j;
  } else
    r = i;
  agg[i] = id[r];
}

__global__ void k_compose(int n, int* __restrict__ agg, const int* __restrict__ agg2) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) agg[i] = agg2[agg[i]];
}

__global__ void k_count_members(int n, const int* __restrict__ agg, int* __restrict__ cnt) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) atomicAdd(&cnt[agg[i]], 1);
}

__global__ void k_fill_members(int n, const int* __restrict__ agg, int* __restrict__ cursor,
                               int* __restrict__ mem) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) mem[atomicAdd(&cursor[agg[i]], 1)] = i;
}

// mpos[mem[k]] = k: each row's place in the member list
__global__ void k_member_pos(int n, const int* __restrict__ mem, int* __restrict__ mpos) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) mpos[mem[k]] = k;
}

// N values in registers into ascending order (odd-even transposition network)
template <int N>
__device__ __forceinline__ void sort_regs(int (&v)[N]) {
#pragma unroll
  for (int round = 0; round < N; ++round) {
#pragma unroll
    for (int u = round & 1; u + 1 < N; u += 2) {
      const int lo = min(v[u], v[u + 1]), hi = max(v[u], v[u + 1]);
      v[u] = lo;
      v[u + 1] = hi;
    }
  }
}
template <int N>
__device__ __forceinline__ void sort_run(int* __restrict__ a, int len) {
  int v[N];
#pragma unroll
  for (int u = 0; u < N; ++u) v[u] = u < len ? a[u] : 0x7fffffff;
  sort_regs<N>(v);
#pragma unroll
  for (int u = 0; u < N; ++u)
    if (u < len) a[u] = v[u];
}

// members of each aggregate into ascending order: up to 16 or 32 sorted in
// registers (padding INT_MAX), larger aggregates by insertion in place
// (composed pairwise aggregates hold about 16-24 members)
__global__ void k_sort_members(int nc, const int* __restrict__ moff, int* __restrict__ mem) {
  int I = blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= nc) return;
  int b = moff[I], e = moff[I + 1];
  if (e - b <= 16) {
    sort_run<16>(mem + b, e - b);
    return;
  }
  if (e - b <= 32) {
    sort_run<32>(mem + b, e - b);
    return;
  }
  for (int a = b + 1; a < e; ++a) {
    int x = mem[a];
    int j = a - 1;
    while (j >= b && mem[j] > x) {
      mem[j + 1] = mem[j];
      --j;
    }
    mem[j + 1] = x;
  }
}

// ---------------------------------------------------------- Galerkin ---
// Aggregate sum C = P^T A P, one 64-lane wave per coarse row I.  The fine
// entries of I's members are visited in member order and CSR order within
// a row; the members' rows are laid end to end (offsets from a wave scan of
// their lengths), so the wave takes 64 entries at a time across members
// (a level-0 row has ~7 entries: one member per step left most lanes idle
// and paid the dependent loads once per member).  The distinct coarse
// columns are kept in an LDS list in first-seen order, and a column's value
// is its first entry plus the later ones in visiting order, whatever the
// chunking (the serial definition's order, so the sums are bit-identical).
// FILL 0 counts the distinct columns (more than kGalCap: overflow); FILL 1
// writes them sorted by column.
static constexpr int kGalWaves = 4;
template <bool FILL>
__global__ __launch_bounds__(64 * kGalWaves) void k_gal_wave(
    int nc, const int* __restrict__ moff, const int* __restrict__ mem, const int* __restrict__ ro,
    const int* __restrict__ ci, const double* __restrict__ va, const int* __restrict__ agg,
    int* __restrict__ cnt, int* __restrict__ overflow, const int* __restrict__ cro,
    int* __restrict__ cci, double* __restrict__ cva) {
  __shared__ int sJ[kGalWaves][kGalCap];
  __shared__ double sV[kGalWaves][kGalCap];
  __shared__ int eJ[kGalWaves][64], eT[kGalWaves][64];
  __shared__ double eV[kGalWaves][64];
  __shared__ int mS[kGalWaves][64], mO[kGalWaves][64], mL[kGalWaves][64];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int I = blockIdx.x * kGalWaves + w;
  if (I >= nc) return;  // the whole wave: no block barriers below
  int* lJ = sJ[w];
  double* lV = sV[w];
  int nd = 0;
  bool over = false;
  const int m1 = moff[I + 1];
  for (int mb = moff[I]; mb < m1 && !over; mb += 64) {  // members, 64 at a time
    const int nm = min(64, m1 - mb);
    int rs = 0, rl = 0;
    if (lane < nm) {
      const int i = mem[mb + lane];
      rs = ro[i];
      rl = ro[i + 1] - rs;
    }
    int inc = rl;  // inclusive scan of the row lengths over the lanes
    for (int off = 1; off < 64; off <<= 1) {
      const int t = __shfl_up(inc, off, 64);
      if (lane >= off) inc += t;
    }
    const int total = __shfl(inc, 63, 64);
    mS[w][lane] = rs;
    mO[w][lane] = inc - rl;
    mL[w][lane] = rl;
    wave_sync();
    for (int k0 = 0; k0 < total; k0 += 64) {
      const int e = k0 + lane;  // position in the members' entries laid end to end
      const bool has = e < total;
      const int ke = total;     // chunk bound for the lead-lane sums below
      int J = -1;
      double v = 0.0;
      if (has) {
        int q = 0;
        for (int p = 0; p < nm; ++p)
          if (mO[w][p] <= e && e < mO[w][p] + mL[w][p]) q = p;
        const int k = mS[w][q] + (e - mO[w][q]);
        J = agg[ci[k]];
        if (FILL) v = va[k];
      }
      eJ[w][lane] = J;
      if (FILL) eV[w][lane] = v;
      wave_sync();
      int t = -1;
      if (has)
        for (int q = 0; q < nd; ++q)
          if (lJ[q] == J) {
            t = q;
            break;
          }
      bool first = has && t < 0;
      if (first)
        for (int l = 0; l < lane; ++l)
          if (eJ[w][l] == J) {
            first = false;
            break;
          }
      const unsigned long long newMask = __ballot(first);
      const int nnew = __popcll(newMask);
      if (nd + nnew > kGalCap) {
        over = true;
        break;
      }
      if (first) {
        t = nd + __popcll(newMask & ((1ull << lane) - 1ull));
        lJ[t] = J;
      }
      wave_sync();
      if (has && t < 0)  // a later repeat of a column new in this chunk
        for (int q = nd; q < nd + nnew; ++q)
          if (lJ[q] == J) {
            t = q;
            break;
          }
      const int ndOld = nd;
      nd += nnew;
      if (FILL) {
        eT[w][lane] = t;
        wave_sync();
        // the lowest lane of each target adds the chunk's entries of that
        // target in lane order onto the running value
        bool lead = has;
        for (int l = 0; l < lane && lead; ++l)
          if (eT[w][l] == t) lead = false;
        if (lead) {
          const int nE = min(64, ke - k0);  // entries in this chunk
          double acc = (t >= ndOld) ? v : lV[t] + v;
          for (int l = lane + 1; l < nE; ++l)
            if (eT[w][l] == t) acc += eV[w][l];
          lV[t] = acc;
        }
      }
      wave_sync();
    }
  }
  if (!FILL) {
    if (lane == 0) {
      cnt[I] = nd;
      if (over) atomicOr(overflow, 1);
    }
    return;
  }
  const int o = cro[I];
  for (int q = lane; q < nd; q += 64) {
    const int J = lJ[q];
    int rank = 0;
    for (int p = 0; p < nd; ++p) rank += lJ[p] < J ? 1 : 0;
    cci[o + rank] = J;
    cva[o + rank] = lV[q];
  }
}

template <typename T>
__global__ void k_omdinv(int n, const int* __restrict__ ro, const int* __restrict__ ci,
                         const double* __restrict__ va, double omega, T* __restrict__ out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double d = 0.0;
  for (int k = ro[i]; k < ro[i + 1]; ++k)
    if (ci[k] == i) d = va[k];
  out[i] = (T)((d != 0.0) ? omega * (1.0 / d) : 0.0);
}

// omdinv_i = cc[0] / a_ii written as k_omdinv writes omega / a_ii
template <typename T>
__global__ void k_omdinv_dev(int n, const int* __restrict__ ro, const int* __restrict__ ci,
                             const double* __restrict__ va, const double* __restrict__ cc,
                             T* __restrict__ out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double d = 0.0;
  for (int k = ro[i]; k < ro[i + 1]; ++k)
    if (ci[k] == i) d = va[k];
  out[i] = (T)((d != 0.0) ? cc[0] * (1.0 / d) : 0.0);
}

// Block maxima of the Gershgorin bound of D^-1 A: sum_k |a_ik| / |a_ii| over
// rows with a_ii != 0 (each row sum in row order; a maximum is the same
// number in any order of evaluation)
__global__ __launch_bounds__(kB) void k_gersh(int n, const int* __restrict__ ro,
                                              const int* __restrict__ ci,
                                              const double* __restrict__ va,
                                              double* __restrict__ partial) {
  __shared__ double sh[kW];
  double g = 0.0;
  for (int i = blockIdx.x * kB + threadIdx.x; i < n; i += gridDim.x * kB) {
    double s = 0.0, d = 0.0;
    for (int k = ro[i]; k < ro[i + 1]; ++k) {
      s += fabs(va[k]);
      if (ci[k] == i) d = va[k];
    }
    if (d != 0.0) g = fmax(g, s / fabs(d));
  }
  for (int off = 32; off > 0; off >>= 1) g = fmax(g, __shfl_down(g, off, 64));
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = g;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < kW; ++k) g = fmax(g, sh[k]);
    partial[blockIdx.x] = g;
  }
}

// The Chebyshev coefficients from the bound (the oracle's cheb_coefficients,
// expression for expression): lmin = fraction lmax, theta, delta, sigma,
// rho1 = 1/sigma, rho2 = 1/(2 sigma - rho1); cc = {1/theta, 1 + rho2 rho1,
// -(rho2 rho1), (2 rho2 / delta) theta}
__global__ __launch_bounds__(kB) void k_cheb_coef(int nparts, const double* __restrict__ partial,
                                                  double fraction, double* __restrict__ cc) {
  __shared__ double sh[kW];
  double g = 0.0;
  for (int i = threadIdx.x; i < nparts; i += kB) g = fmax(g, partial[i]);
  for (int off = 32; off > 0; off >>= 1) g = fmax(g, __shfl_down(g, off, 64));
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = g;
  __syncthreads();
  if (threadIdx.x != 0) return;
  for (int k = 1; k < kW; ++k) g = fmax(g, sh[k]);
  const double lmax = g > 0.0 ? g : 1.0;
  const double lmin = fraction * lmax;
  const double theta = 0.5 * (lmax + lmin), delta = 0.5 * (lmax - lmin);
  const double sigma = theta / delta;
  const double rho1 = 1.0 / sigma;
  const double rho2 = 1.0 / (2.0 * sigma - rho1);
  cc[0] = 1.0 / theta;
  cc[1] = 1.0 + rho2 * rho1;
  cc[2] = -(rho2 * rho1);
  cc[3] = (2.0 * rho2 / delta) * theta;
}

template <typename T>
__global__ void k_cast(long long n, const double* __restrict__ a, T* __restrict__ b) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (T)a[i];
}

// Coefficients of the two flexible-CG steps from ks = {c1.Ac1, c1.b,
// c2.v1, c2.Ac2, c2.b}; a zero c1.Ac1 (beta <= 0) drops the terms it divides.
struct KCoef {
  double a1, gr, a2;
};
__device__ __forceinline__ double kcyc_a1(const double* ks) {
  return ks[0] > 0.0 ? ks[1] / ks[0] : 0.0;
}
__device__ __forceinline__ KCoef kcyc_coef(const double* ks) {
  KCoef k;
  k.a1 = kcyc_a1(ks);
  k.gr = ks[0] > 0.0 ? ks[2] / ks[0] : 0.0;
  const double beta = ks[3] - ks[2] * k.gr;
  k.a2 = beta > 0.0 ? ks[4] / beta : 0.0;
  return k;
}

// --------------------------------------------------------- V-cycle -----
// One level of the V-cycle (omega-damped Jacobi, one sweep each way, from
// x = 0), with y = omdinv o b supplied by the producer of b:
//   down:     r_i = b_i - sum_j a_ij y_j               (KIND 0, v = y)
//   restrict: bc_I = sum_{i in I} r_i, yc_I = omdinv_c,I bc_I
//   prolong:  e_i = y_i + xc[agg_i]
//   up:       x_i = e_i + omdinv_i (b_i - sum_j a_ij e_j)   (KIND 1, v = e)
// so the two CSR-stream kernels gather a single vector, like an SpMV.
// The degree-2 Chebyshev smoother's second step (cc = {1/theta, cv, cw, cr},
// omdinv = (1/theta) / a_ii):
//   x_i = cv v_i + cr omdinv_i (b_i - sum_j a_ij v_j)          (KIND 2, from 0)
//   x_i = (cv v_i + cw wv_i) + cr omdinv_i (b_i - sum_j a_ij v_j) (KIND 3)
// DOT (KIND 1 and 3) adds
// the partial sums of b_i x_i (DOT 1) and x_i qv_i (DOT 2) of the result and
// finishes them in the kernel: the preconditioned CG's r.z and z.q.
// KIND 0 with opos stores r_i at opos[i], its place in the aggregation's
// member list, so the restriction sums contiguous runs.
// C16: columns from the 16-bit offsets c16 (ColOffsets) instead of ci.
template <typename T, int KIND, int DOT, bool C16>
__global__ __launch_bounds__(kB) void k_level_rows(int nrb, const int2* __restrict__ rbs,
                                                   const int* __restrict__ ro,
                                                   const int* __restrict__ ci,
                                                   const int16_t* __restrict__ c16,
                                                   const T* __restrict__ va,
                                                   const T* __restrict__ omdinv,
                                                   const T* __restrict__ b,
                                                   const T* __restrict__ v,
                                                   T* __restrict__ out,
                                                   const T* __restrict__ qv, RedArgs ra,
                                                   const T* __restrict__ wv = nullptr,
                                                   const double* __restrict__ cc = nullptr,
                                                   const int* __restrict__ opos = nullptr) {
  __shared__ double prod[kW][kCap];
  __shared__ double sh[2 * kW];
  const int need = (nrb + kW - 1) / kW;
  int lb;
  if ((int)gridDim.x == 8 * need) {
    lb = spread_block(need);
    if (lb < 0) return;  // whole block: no barrier is left waiting
  } else {
    lb = xcd_block(gridDim.x);
  }
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const WaveRows q = wave_rows(nrb, rbs, lb * kW + w);
  const int i = q.r0 + lane;
  auto vj = [&](int j) { return (double)v[j]; };
  int kb = 0, ke = 0, oi_pos = i;
  double bi = 0.0, vi = 0.0, oi = 0.0, qi = 0.0, wi = 0.0;
  if (i < q.r1) {
    kb = ro[i];
    ke = ro[i + 1];
    bi = (double)b[i];
    if (KIND == 0 && opos) oi_pos = opos[i];
    if (KIND >= 1) {
      vi = (double)v[i];
      oi = (double)omdinv[i];
    }
    if (KIND == 3) wi = (double)wv[i];
    if (DOT == 2) qi = (double)qv[i];
  }
  if (q.stream && q.k1 > q.k0) {
    if constexpr (C16)
      stage_products(q, lane, ColOffsets{c16, q.r0}, va, vj, prod[w]);
    else
      stage_products(q, lane, ci, va, vj, prod[w]);
  }
  wave_sync();
  double d[2] = {0.0, 0.0};
  if (i < q.r1) {
    double s = 0.0;
    if (q.stream) {
      for (int k = kb; k < ke; ++k) s += prod[w][k - q.k0];
    } else {
      for (int k = kb; k < ke; ++k) s += (double)va[k] * vj(C16 ? q.r0 + (int)c16[k] : ci[k]);
    }
    if (KIND == 0) {
      out[oi_pos] = (T)(bi - s);
    } else {
      T o;
      if (KIND == 1) {
        o = (T)(vi + oi * (bi - s));
      } else {
        const double t = oi * (bi - s);
        o = KIND == 2 ? (T)(cc[1] * vi + cc[3] * t) : (T)((cc[1] * vi + cc[2] * wi) + cc[3] * t);
      }
      out[i] = o;
      if (DOT >= 1) d[0] = bi * (double)o;
      if (DOT == 2) d[1] = (double)o * qi;
    }
  }
  if constexpr (DOT > 0) {
    block_sums<DOT>(d, sh);
    store_partials<DOT>(d, lb, ra);
  }
}

// r.z (and z.q) of a preconditioner application that ended on the coarsest
// level (no up kernel to carry them)
template <typename T, int DOT>
__global__ __launch_bounds__(kB) void k_apply_dots(int n, const T* __restrict__ r,
                                                   const T* __restrict__ z,
                                                   const T* __restrict__ qv, RedArgs ra) {
  __shared__ double sh[2 * kW];
  double d[2] = {0.0, 0.0};
  for (int i = blockIdx.x * kB + threadIdx.x; i < n; i += gridDim.x * kB) {
    const double zi = (double)z[i];
    d[0] += (double)r[i] * zi;
    if (DOT == 2) d[1] += zi * (double)qv[i];
  }
  block_sums<DOT>(d, sh);
  store_partials<DOT>(d, blockIdx.x, ra);
}

// y = omdinv o b (level-0 entry of the V-cycle)
template <typename T>
__global__ void k_scale(int n, const T* __restrict__ omdinv, const T* __restrict__ b,
                        T* __restrict__ y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = (T)((double)omdinv[i] * (double)b[i]);
}

// bc_I = sum of the members' residuals in ascending member order;
// yc_I = omdinv_c,I * bc_I for the coarse level's own down kernel.  The
// residual is read as r[mem[k]] or, with mem null, from r stored in
// member-list order (r[k] is member k's residual, as the down sweep writes
// it on coarse levels).  One thread per aggregate, loads in batches of 8
// (all index loads, then all values; slots past the end re-read the last
// member and are not added).
template <typename T>
__global__ void k_restrict(int nc, const int* __restrict__ moff, const int* __restrict__ mem,
                           const T* __restrict__ r, const T* __restrict__ omdc,
                           T* __restrict__ bc, T* __restrict__ yc) {
  constexpr int U = 8;
  const int lb = spread_block((nc + (int)blockDim.x - 1) / (int)blockDim.x);
  if (lb < 0) return;
  const int I = lb * blockDim.x + threadIdx.x;
  if (I >= nc) return;
  const int m0 = moff[I], m1 = moff[I + 1];  // m1 > m0: aggregates are not empty
  double s = 0.0;
  for (int b = m0; b < m1; b += U) {
    int idx[U];
    double v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) idx[u] = mem ? mem[min(b + u, m1 - 1)] : min(b + u, m1 - 1);
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = (double)r[idx[u]];
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (b + u < m1) s += v[u];
  }
  const T bI = (T)s;
  bc[I] = bI;
  yc[I] = (T)((double)omdc[I] * (double)bI);
}

// The same aggregate sums over row blocks of the member lists (moff as row
// offsets, mem as columns; one wave per block): the lanes first load all of
// the block's members' r values (with mem: coalesced member reads, one
// gather round trip; without: unit-stride loads), stage them in the wave's
// LDS slice, then lane I adds aggregate I's values in member order --
// k_restrict's order, so the same bits, without one thread walking a whole
// aggregate.
template <typename T>
__global__ __launch_bounds__(kB) void k_restrict_rb(int nrb, const int2* __restrict__ rbs,
                                                    const int* __restrict__ moff,
                                                    const int* __restrict__ mem,
                                                    const T* __restrict__ r,
                                                    const T* __restrict__ omdc,
                                                    T* __restrict__ bc, T* __restrict__ yc) {
  __shared__ double vals[kW][kCap];
  const int lb = spread_block((nrb + kW - 1) / kW);
  if (lb < 0) return;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const WaveRows q = wave_rows(nrb, rbs, lb * kW + w);
  const int I = q.r0 + lane;
  int kb = 0, ke = 0;
  if (I < q.r1) {
    kb = moff[I];
    ke = moff[I + 1];
  }
  if (q.stream && q.k1 > q.k0) {
    int m[kStreamPer];
    double v[kStreamPer];
#pragma unroll
    for (int u = 0; u < kStreamPer; ++u) {
      const int k = min(q.k0 + lane + 64 * u, q.k1 - 1);
      m[u] = mem ? mem[k] : k;
    }
#pragma unroll
    for (int u = 0; u < kStreamPer; ++u) v[u] = (double)r[m[u]];
#pragma unroll
    for (int u = 0; u < kStreamPer; ++u) {
      const int k = q.k0 + lane + 64 * u;
      if (k < q.k1) vals[w][k - q.k0] = v[u];
    }
  }
  wave_sync();
  if (I >= q.r1) return;
  double s = 0.0;
  if (q.stream)
    for (int k = kb; k < ke; ++k) s += vals[w][k - q.k0];
  else
    for (int k = kb; k < ke; ++k) s += (double)r[mem ? mem[k] : k];
  const T bI = (T)s;
  bc[I] = bI;
  yc[I] = (T)((double)omdc[I] * (double)bI);
}

// e_i = y_i + xc[agg_i]  (e may be y: an in-place, per-element update).
// KC: the coarse level took the K-cycle correction; xc holds its second
// step c2 and the correction (T)((a1 - a2 g/rho1) c1 + a2 c2) is formed
// here from the step's scalars ks.  (Two or four elements per thread
// measured slower at 2.08 M rows.)
template <typename T, bool KC>
__global__ void k_prolong(int n, const T* y, const T* __restrict__ xc, const T* __restrict__ c1,
                          const double* __restrict__ ks, const int* __restrict__ agg, T* e) {
  const int lb = spread_block((n + (int)blockDim.x - 1) / (int)blockDim.x);
  if (lb < 0) return;
  const int i = lb * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int a = agg[i];
  double xcv;
  if (KC) {
    const KCoef k = kcyc_coef(ks);
    const double al1 = k.a1 - k.a2 * k.gr;
    xcv = (double)(T)(al1 * (double)c1[a] + k.a2 * (double)xc[a]);
  } else {
    xcv = (double)xc[a];
  }
  e[i] = (T)((double)y[i] + xcv);
}

// coarsest: x_i = sum_j Ainv[i][j] b_j, Ainv stored column-major (FP64).
// Blocked order: lane = row, wave = column chunk of CH = 16 (n <= 1024) or
// 64 (n <= 4096) columns; each chunk's partial is summed in column order,
// then the partials in chunk order.  A chunk's loads are all issued before
// the sum, so a wave waits on one memory round trip per chunk.
static constexpr int kDenseThreads = 1024;
static inline int dense_chunk(int n) { return n <= 1024 ? 16 : 64; }
template <typename T, int CH>
__global__ __launch_bounds__(kDenseThreads) void k_dense_solve(int n,
                                                               const double* __restrict__ invT,
                                                               const T* __restrict__ b,
                                                               T* __restrict__ x) {
  __shared__ double part[64][64];
  const int lb = spread_block((n + 63) / 64);
  if (lb < 0) return;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int nw = blockDim.x >> 6;
  const int i = lb * 64 + lane;
  const int ic = min(i, n - 1);
  const int nch = (n + CH - 1) / CH;
  for (int c = w; c < nch; c += nw) {
    double t = 0.0;
    for (int j0 = c * CH; j0 < c * CH + CH; j0 += 16) {  // batches of 16 loads, summed in order
      double a[16], bb[16];
#pragma unroll
      for (int u = 0; u < 16; ++u) {
        const int j = min(j0 + u, n - 1);
        a[u] = invT[(size_t)j * n + ic];
        bb[u] = (double)b[j];
      }
#pragma unroll
      for (int u = 0; u < 16; ++u)
        if (j0 + u < n) t += a[u] * bb[u];
    }
    part[c][lane] = t;
  }
  __syncthreads();
  if (w == 0 && i < n) {
    double s = 0.0;
    for (int c = 0; c < nch; ++c) s += part[c][lane];
    x[i] = (T)s;
  }
}

// plain damped Jacobi sweep (coarsest fallback when too large for dense)
template <typename T>
__global__ void k_jacobi(int n, const int* __restrict__ ro, const int* __restrict__ ci,
                         const T* __restrict__ va, const T* __restrict__ omdinv,
                         const T* __restrict__ b, const T* __restrict__ x, T* __restrict__ y) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int k = ro[i]; k < ro[i + 1]; ++k) s += (double)va[k] * (double)x[ci[k]];
  y[i] = (T)((double)x[i] + (double)omdinv[i] * ((double)b[i] - s));
}

// K-cycle products on a coarse level (a CSR-stream pass gathering c):
//   NC 2: w = A c stored in w; partial sums of c.w and c.b
//   NC 3: w = A c not stored; partial sums of c.v1, c.w and c.b
// The products use w rounded to T, the value an FP32 cycle stores.
template <typename T, int NC>
__global__ __launch_bounds__(kB) void k_kprod(int nrb, const int2* __restrict__ rbs,
                                              const int* __restrict__ ro,
                                              const int* __restrict__ ci,
                                              const T* __restrict__ va, const T* __restrict__ c,
                                              const T* __restrict__ b, const T* __restrict__ v1,
                                              T* __restrict__ w, RedArgs ra) {
  __shared__ double prod[kW][kCap];
  __shared__ double sh[NC * kW];
  const int lb = xcd_block(gridDim.x);
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const WaveRows q = wave_rows(nrb, rbs, lb * kW + wv);
  const int i = q.r0 + lane;
  int kb = 0, ke = 0;
  double d[NC];
#pragma unroll
  for (int k = 0; k < NC; ++k) d[k] = 0.0;
  if (i < q.r1) {
    kb = ro[i];
    ke = ro[i + 1];
  }
  if (q.stream && q.k1 > q.k0)
    stage_products(q, lane, ci, va, [&](int j) { return (double)c[j]; }, prod[wv]);
  wave_sync();
  if (i < q.r1) {
    double s = 0.0;
    if (q.stream) {
      for (int k = kb; k < ke; ++k) s += prod[wv][k - q.k0];
    } else {
      for (int k = kb; k < ke; ++k) s += (double)va[k] * (double)c[ci[k]];
    }
    const T wi = (T)s;
    const double cv = (double)c[i];
    if (NC == 2) {
      w[i] = wi;
      d[0] = cv * (double)wi;
      d[1] = cv * (double)b[i];
    } else {
      d[0] = cv * (double)v1[i];
      d[1] = cv * (double)wi;
      d[NC - 1] = cv * (double)b[i];
    }
  }
  block_sums<NC>(d, sh);
  store_partials<NC>(d, lb, ra);
}

// b -= a1 v1 (the second step's right-hand side, in place); y = omdinv o b.
// Every block first finishes ks[0..1] = {c1.Ac1, c1.b} from k_kprod<T, 2>'s
// nparts block partials itself, in the order of the 1024-thread
// k_finalize_nc<2> (its 1024 threads' strided sums, each wave's shuffle
// tree, then the 16 wave sums in order), so the scalars have the bits a
// separate finalize would give; block 0 stores them for the prolongation.
// A coarse level has few blocks, so the repeated sums cost less than the
// separate launch.
static constexpr int kFinThreads = 1024;  // the finalize order reproduced
template <typename T>
__global__ __launch_bounds__(kB) void k_kres(int n, int nparts, const double* __restrict__ partial,
                                             double* __restrict__ ks, const T* __restrict__ v1,
                                             const T* __restrict__ omd, T* __restrict__ b,
                                             T* __restrict__ y) {
  constexpr int J = kFinThreads / kB;  // finalize threads per thread here
  constexpr int NW = kFinThreads / 64;
  __shared__ double sh[2 * NW];
  __shared__ double a1s;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  double acc[J][2];
#pragma unroll
  for (int j = 0; j < J; ++j) {
    acc[j][0] = acc[j][1] = 0.0;
    for (int i = threadIdx.x + kB * j; i < nparts; i += kFinThreads) {
      acc[j][0] += partial[2 * (size_t)i];
      acc[j][1] += partial[2 * (size_t)i + 1];
    }
  }
#pragma unroll
  for (int j = 0; j < J; ++j)
#pragma unroll
    for (int c = 0; c < 2; ++c)
      for (int off = 32; off > 0; off >>= 1) acc[j][c] += __shfl_down(acc[j][c], off, 64);
  if (lane == 0)
#pragma unroll
    for (int j = 0; j < J; ++j)
#pragma unroll
      for (int c = 0; c < 2; ++c) sh[c * NW + w + (kB / 64) * j] = acc[j][c];
  __syncthreads();
  if (threadIdx.x == 0) {
    double k2[2];
    for (int c = 0; c < 2; ++c) {
      double t = 0.0;
      for (int i = 0; i < NW; ++i) t += sh[c * NW + i];
      k2[c] = t;
    }
    if (blockIdx.x == 0) {
      ks[0] = k2[0];
      ks[1] = k2[1];
    }
    a1s = kcyc_a1(k2);
  }
  __syncthreads();
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double a1 = a1s;
  const T bi = (T)((double)b[i] - a1 * (double)v1[i]);
  b[i] = bi;
  y[i] = (T)((double)omd[i] * (double)bi);
}


// --------------------------------------------------- dense coarsest ----
// A[i][:] = sum of row i's CSR entries in CSR order (duplicates added in
// order), |A[i][i]| into dabs[i]; one thread per row.
__global__ void k_dense_scatter(int n, const int* __restrict__ ro, const int* __restrict__ ci,
                                const double* __restrict__ va, double* __restrict__ A,
                                double* __restrict__ dabs) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double* row = A + (size_t)i * n;
  for (int j = 0; j < n; ++j) row[j] = 0.0;
  for (int k = ro[i]; k < ro[i + 1]; ++k) row[ci[k]] += va[k];
  dabs[i] = fabs(row[i]);
}

// Dense inverse of the coarsest level in three stages, every entry formed
// by the oracle's sequence of operations (dense_inverse, oracle/solvers.c):
//   1. Cholesky A = L L^T, L in the lower triangle of A;
//   2. M = L^-1 by forward substitution (a zero pivot gives a zero row);
//   3. inverse = M^T M, entry (i, j) summed over k >= max(i, j) in
//      increasing k (symmetric bit for bit, so row- and column-major agree).
// Stages 1 and 2 run right-looking in panels of kPanel pivots: a one-
// workgroup kernel walks the panel's pivots, a tiled kernel applies the
// panel to the trailing part.  Right-looking applies to each entry the
// updates j = 0, 1, ... in increasing order, exactly the left-looking sums
// of the oracle, so L and M carry its bits.  A pivot d <= 1e-13 max|diag|
// zeroes its column of L.
static constexpr int kPanel = 32;
static constexpr int kPanelThreads = 1024;
static constexpr int kTile = 16;

// thr[0] = 1e-13 max_i |a_ii| (a maximum: order-free)
__global__ __launch_bounds__(kPanelThreads) void k_dense_thr(int n,
                                                             const double* __restrict__ dabs,
                                                             double* __restrict__ thr) {
  __shared__ double red[kPanelThreads / 64];
  const int t = threadIdx.x;
  double m = 0.0;
  for (int i = t; i < n; i += blockDim.x) m = fmax(m, dabs[i]);
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_down(m, off, 64));
  if ((t & 63) == 0) red[t >> 6] = m;
  __syncthreads();
  if (t == 0) {
    double dm = 0.0;
    for (int w = 0; w < (int)(blockDim.x / 64); ++w) dm = fmax(dm, red[w]);
    thr[0] = 1e-13 * dm;
  }
}

// pivots k = p0 .. p1-1: the pivot, column k of L (rows below), then the
// update of the panel's later columns j (rows i >= j) by column k
__global__ __launch_bounds__(kPanelThreads) void k_chol_panel(int n, int p0, int p1,
                                                              double* __restrict__ A,
                                                              const double* __restrict__ thr) {
  __shared__ double lk;
  __shared__ int skip;
  const int t = threadIdx.x, nt = blockDim.x;
  for (int k = p0; k < p1; ++k) {
    if (t == 0) {
      const double d = A[(size_t)k * n + k];
      skip = d <= thr[0];
      if (!skip) {
        const double l = sqrt(d);
        A[(size_t)k * n + k] = l;
        lk = l;
      }
    }
    __syncthreads();
    if (skip) {
      for (int i = k + t; i < n; i += nt) A[(size_t)i * n + k] = 0.0;
    } else {
      const double l = lk;
      for (int i = k + 1 + t; i < n; i += nt) A[(size_t)i * n + k] = A[(size_t)i * n + k] / l;
    }
    __syncthreads();
    const int nj = p1 - k - 1;
    if (nj > 0) {
      const long long ne = (long long)(n - k - 1) * nj;
      for (long long e = t; e < ne; e += nt) {
        const int i = k + 1 + (int)(e / nj), j = k + 1 + (int)(e % nj);
        if (j <= i) A[(size_t)i * n + j] -= A[(size_t)i * n + k] * A[(size_t)j * n + k];
      }
    }
    __syncthreads();
  }
}

// trailing part after the panel [p0, p1): a_ij -= l_ik l_jk, k = p0 .. p1-1
// in order, for j >= p1 and i >= j; one 16 x 16 tile of (i, j) per block
__global__ __launch_bounds__(kTile* kTile) void k_chol_trail(int n, int p0, int p1,
                                                             double* __restrict__ A) {
  __shared__ double Li[kTile][kPanel + 1], Lj[kTile][kPanel + 1];
  const int T = (n - p1 + kTile - 1) / kTile;
  const int ti = blockIdx.x / T, tj = blockIdx.x % T;
  if (ti < tj) return;  // whole block: the lower triangle of tiles only
  const int i0 = p1 + ti * kTile, j0 = p1 + tj * kTile, w = p1 - p0;
  for (int e = threadIdx.x; e < kTile * w; e += blockDim.x) {
    const int r = e / w, k = e % w;
    Li[r][k] = i0 + r < n ? A[(size_t)(i0 + r) * n + p0 + k] : 0.0;
    Lj[r][k] = j0 + r < n ? A[(size_t)(j0 + r) * n + p0 + k] : 0.0;
  }
  __syncthreads();
  const int ri = threadIdx.x / kTile, rj = threadIdx.x % kTile;
  const int i = i0 + ri, j = j0 + rj;
  if (i >= n || j >= n || j > i) return;
  double a = A[(size_t)i * n + j];
  for (int k = 0; k < w; ++k) a -= Li[ri][k] * Lj[rj][k];
  A[(size_t)i * n + j] = a;
}

// forward substitution, pivots k = p0 .. p1-1: row k of M (columns c <= k),
// then the panel's later rows i updated by it
__global__ __launch_bounds__(kPanelThreads) void k_fwd_panel(int n, int p0, int p1,
                                                             const double* __restrict__ L,
                                                             double* __restrict__ S) {
  const int t = threadIdx.x, nt = blockDim.x;
  for (int k = p0; k < p1; ++k) {
    const double l = L[(size_t)k * n + k];
    for (int c = t; c <= k; c += nt) {
      const double v = S[(size_t)k * n + c];
      S[(size_t)k * n + c] = (l != 0.0) ? v / l : 0.0;
    }
    __syncthreads();
    const int ni = p1 - k - 1;
    if (ni > 0) {
      const long long ne = (long long)ni * (k + 1);
      for (long long e = t; e < ne; e += nt) {
        const int i = k + 1 + (int)(e / (k + 1)), c = (int)(e % (k + 1));
        S[(size_t)i * n + c] -= L[(size_t)i * n + k] * S[(size_t)k * n + c];
      }
    }
    __syncthreads();
  }
}

// trailing rows after the panel: s_ic -= l_ik m_kc, k = max(p0, c) .. p1-1 in
// order (the terms k < c are zero products, which the oracle also adds to
// a zero), for i >= p1 and c < p1; one 16 x 16 tile of (i, c) per block
__global__ __launch_bounds__(kTile* kTile) void k_fwd_trail(int n, int p0, int p1,
                                                            const double* __restrict__ L,
                                                            double* __restrict__ S) {
  __shared__ double Lt[kTile][kPanel + 1], Mt[kPanel][kTile + 1];
  const int TC = (p1 + kTile - 1) / kTile;
  const int ti = blockIdx.x / TC, tc = blockIdx.x % TC;
  const int i0 = p1 + ti * kTile, c0 = tc * kTile, w = p1 - p0;
  for (int e = threadIdx.x; e < kTile * w; e += blockDim.x) {
    const int r = e / w, k = e % w;
    Lt[r][k] = i0 + r < n ? L[(size_t)(i0 + r) * n + p0 + k] : 0.0;
    const int kk = e / kTile, cc = e % kTile;
    Mt[kk][cc] = c0 + cc < p1 ? S[(size_t)(p0 + kk) * n + c0 + cc] : 0.0;
  }
  __syncthreads();
  const int ri = threadIdx.x / kTile, rc = threadIdx.x % kTile;
  const int i = i0 + ri, c = c0 + rc;
  if (i >= n || c >= p1) return;
  double v = S[(size_t)i * n + c];
  for (int k = max(c - p0, 0); k < w; ++k) v -= Lt[ri][k] * Mt[k][rc];
  S[(size_t)i * n + c] = v;
}

// inverse = M^T M: entries (i, j) and (j, i) of a 16 x 16 tile pair, the sum
// over k in increasing order from max(i0, j0) (terms below max(i, j) are
// zero products added to zero), M staged 32 rows at a time
__global__ __launch_bounds__(kTile* kTile) void k_mtm(int n, const double* __restrict__ M,
                                                      double* __restrict__ inv) {
  __shared__ double Mi[kPanel][kTile + 1], Mj[kPanel][kTile + 1];
  const int T = (n + kTile - 1) / kTile;
  const int ti = blockIdx.x / T, tj = blockIdx.x % T;
  if (ti < tj) return;  // whole block: tile pairs below the diagonal
  const int i0 = ti * kTile, j0 = tj * kTile;
  const int ri = threadIdx.x / kTile, rj = threadIdx.x % kTile;
  const int i = i0 + ri, j = j0 + rj;
  double s = 0.0;
  for (int kb = i0; kb < n; kb += kPanel) {  // i0 = max(i0, j0)
    for (int e = threadIdx.x; e < kPanel * kTile; e += blockDim.x) {
      const int kk = e / kTile, cc = e % kTile;
      const bool in = kb + kk < n;
      Mi[kk][cc] = in && i0 + cc < n ? M[(size_t)(kb + kk) * n + i0 + cc] : 0.0;
      Mj[kk][cc] = in && j0 + cc < n ? M[(size_t)(kb + kk) * n + j0 + cc] : 0.0;
    }
    __syncthreads();
    const int kn = min(kPanel, n - kb);
    for (int k = 0; k < kn; ++k) s += Mi[k][ri] * Mj[k][rj];
    __syncthreads();
  }
  if (i < n && j < n) {
    inv[(size_t)i * n + j] = s;
    inv[(size_t)j * n + i] = s;
  }
}

__global__ void k_identity(int n, double* __restrict__ S) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)n * n) return;
  S[e] = (e / (size_t)n == e % (size_t)n) ? 1.0 : 0.0;
}

void dense_inverse_device(const DevCSR<double>& A, DevBuf<double>& invT, hipStream_t s) {
  const int n = A.n;
  if (n <= 0 || n > 4096) throw Error(7, "dense coarsest level must have 1..4096 rows");
  const size_t nn = (size_t)n * n;
  DevBuf<double> L(nn), S(nn), dabs((size_t)n), thr(1);
  k_dense_scatter<<<grid_for(n, 64), 64, 0, s>>>(n, A.rowOff.p, A.cols.p, A.vals.p, L.p, dabs.p);
  k_dense_thr<<<1, kPanelThreads, 0, s>>>(n, dabs.p, thr.p);
  AN_CHECK_LAUNCH();
  for (int p0 = 0; p0 < n; p0 += kPanel) {
    const int p1 = std::min(p0 + kPanel, n);
    k_chol_panel<<<1, kPanelThreads, 0, s>>>(n, p0, p1, L.p, thr.p);
    if (p1 < n) {
      const int T = (n - p1 + kTile - 1) / kTile;
      k_chol_trail<<<T * T, kTile * kTile, 0, s>>>(n, p0, p1, L.p);
    }
    AN_CHECK_LAUNCH();
  }
  k_identity<<<grid_for((long long)nn, 256), 256, 0, s>>>(n, S.p);
  for (int p0 = 0; p0 < n; p0 += kPanel) {
    const int p1 = std::min(p0 + kPanel, n);
    k_fwd_panel<<<1, kPanelThreads, 0, s>>>(n, p0, p1, L.p, S.p);
    if (p1 < n) {
      const int TI = (n - p1 + kTile - 1) / kTile, TC = (p1 + kTile - 1) / kTile;
      k_fwd_trail<<<TI * TC, kTile * kTile, 0, s>>>(n, p0, p1, L.p, S.p);
    }
    AN_CHECK_LAUNCH();
  }
  invT.alloc(nn);
  const int T = (n + kTile - 1) / kTile;
  k_mtm<<<T * T, kTile * kTile, 0, s>>>(n, S.p, invT.p);
  AN_CHECK_LAUNCH();
}

static int read_int(const int* d, hipStream_t s) {
  int h = 0;
  AN_HIP(hipMemcpyAsync(&h, d, sizeof(int), hipMemcpyDeviceToHost, s));
  AN_HIP(hipStreamSynchronize(s));
  return h;
}

// One pairwise pass on W (n rows): agg[n] and the aggregate count.
static int pair_pass(const DevCSR<double>& W, double theta, DevBuf<int>& agg, hipStream_t s) {
  const int n = W.n;
  const int G = grid_for(n, kB);
  DevBuf<double> smax((size_t)n);
  DevBuf<int> match((size_t)n), cand((size_t)n), rep((size_t)n + 1), id((size_t)n + 1);
  k_smax<<<G, kB, 0, s>>>(n, W.rowOff.p, W.cols.p, W.vals.p, smax.p);
  k_fill_int<<<G, kB, 0, s>>>(n, match.p, -1);
  for (int r = 0; r < 6; ++r) {
    k_cand<<<G, kB, 0, s>>>(n, W.rowOff.p, W.cols.p, W.vals.p, smax.p, theta, match.p, cand.p, 0);
    k_match<<<G, kB, 0, s>>>(n, cand.p, match.p);
  }
  k_cand<<<G, kB, 0, s>>>(n, W.rowOff.p, W.cols.p, W.vals.p, smax.p, theta, match.p, cand.p, 1);
  k_rep<<<G, kB, 0, s>>>(n, match.p, cand.p, rep.p);
  AN_CHECK_LAUNCH();
  long long nc = exclusive_scan(rep.p, id.p, n, s);
  agg.alloc((size_t)n);
  k_assign<<<G, kB, 0, s>>>(n, match.p, cand.p, id.p, agg.p);
  AN_CHECK_LAUNCH();
  return (int)nc;
}

static void build_members(int n, int nc, const int* agg, DevBuf<int>& moff, DevBuf<int>& mem,
                          hipStream_t s) {
  DevBuf<int> cnt((size_t)nc + 1);
  cnt.zero(s);
  k_count_members<<<grid_for(n, kB), kB, 0, s>>>(n, agg, cnt.p);
  AN_CHECK_LAUNCH();
  moff.alloc((size_t)nc + 1);
  exclusive_scan(cnt.p, moff.p, nc, s);
  DevBuf<int> cursor((size_t)nc + 1);
  AN_HIP(hipMemcpyAsync(cursor.p, moff.p, sizeof(int) * ((size_t)nc + 1), hipMemcpyDeviceToDevice,
                        s));
  mem.alloc((size_t)n);
  k_fill_members<<<grid_for(n, kB), kB, 0, s>>>(n, agg, cursor.p, mem.p);
  k_sort_members<<<grid_for(nc, kB), kB, 0, s>>>(nc, moff.p, mem.p);
  AN_CHECK_LAUNCH();
}

// Galerkin aggregate sum; returns false (C untouched) when a coarse row would
// hold more than kGalCap entries -- the hierarchy then stops at this level.
static bool galerkin(const DevCSR<double>& W, int nc, const int* agg, const int* moff,
                     const int* mem, DevCSR<double>& C, hipStream_t s) {
  DevBuf<int> cnt((size_t)nc + 1), overflow(1);
  overflow.zero(s);
  const int gw = (nc + kGalWaves - 1) / kGalWaves;
  k_gal_wave<false><<<gw, 64 * kGalWaves, 0, s>>>(nc, moff, mem, W.rowOff.p, W.cols.p, nullptr,
                                                  agg, cnt.p, overflow.p, nullptr, nullptr,
                                                  nullptr);
  AN_CHECK_LAUNCH();
  if (read_int(overflow.p, s)) return false;
  C.n = nc;
  C.rowOff.alloc((size_t)nc + 1);
  C.nnz = exclusive_scan(cnt.p, C.rowOff.p, nc, s);
  C.cols.alloc((size_t)C.nnz);
  C.vals.alloc((size_t)C.nnz);
  k_gal_wave<true><<<gw, 64 * kGalWaves, 0, s>>>(nc, moff, mem, W.rowOff.p, W.cols.p, W.vals.p,
                                                 agg, nullptr, nullptr, C.rowOff.p, C.cols.p,
                                                 C.vals.p);
  AN_CHECK_LAUNCH();
  return true;
}

void AmgHierarchy::setup(const DevCSR<double>& fine, hipStream_t s) {
  PoolScope scope(&pool);
  lv.clear();
  has32 = false;
  lv.reserve(opt.maxLevels + 1);
  lv.emplace_back();
  lv[0].n = fine.n;
  lv[0].A = &fine;
  const int maxL = opt.maxLevels > 1 ? opt.maxLevels : 2;
  for (int l = 0;; ++l) {
    AmgLevel& L = lv[l];
    L.omdinv.alloc((size_t)L.n);
    k_omdinv<double><<<grid_for(L.n, kB), kB, 0, s>>>(L.n, L.A->rowOff.p, L.A->cols.p,
                                                      L.A->vals.p, opt.omega, L.omdinv.p);
    AN_CHECK_LAUNCH();
    L.b.alloc((size_t)L.n);
    L.x.alloc((size_t)L.n);
    L.r.alloc((size_t)L.n);
    L.y.alloc((size_t)L.n);
    if (L.n <= opt.coarseSize || l == maxL - 1) {
      L.isCoarsest = true;
      break;
    }
    // composed pairwise passes
    DevBuf<int> agg;
    int cur = pair_pass(*L.A, opt.strength, agg, s);
    DevCSR<double> W;
    bool fits;
    {
      DevBuf<int> mo, me;
      build_members(L.n, cur, agg.p, mo, me, s);
      fits = galerkin(*L.A, cur, agg.p, mo.p, me.p, W, s);
    }
    for (int p = 1; fits && p < (opt.pairPasses > 0 ? opt.pairPasses : 1); ++p) {
      DevBuf<int> a2;
      int n2 = pair_pass(W, opt.strength, a2, s);
      DevCSR<double> W2;
      DevBuf<int> mo, me;
      build_members(W.n, n2, a2.p, mo, me, s);
      fits = galerkin(W, n2, a2.p, mo.p, me.p, W2, s);
      if (!fits) break;
      k_compose<<<grid_for(L.n, kB), kB, 0, s>>>(L.n, agg.p, a2.p);
      AN_CHECK_LAUNCH();
      W = std::move(W2);
      cur = n2;
    }
    // a coarse operator too dense for the aggregate-sum kernel, or too
    // little coarsening: this level is the coarsest
    if (!fits || cur >= L.n || (double)cur > 0.9 * (double)L.n) {
      L.isCoarsest = true;
      break;
    }
    L.nc = cur;
    L.agg = std::move(agg);
    build_members(L.n, cur, L.agg.p, L.moff, L.mem, s);
    L.mpos.alloc((size_t)L.n);
    k_member_pos<<<grid_for(L.n, kB), kB, 0, s>>>(L.n, L.mem.p, L.mpos.p);
    AN_CHECK_LAUNCH();
    build_row_blocks(cur, L.moff.p, L.mrb, s);
    lv.emplace_back();
    AmgLevel& C = lv.back();
    C.n = cur;
    C.own = std::move(W);
    C.own.finish_pattern(s);
    C.A = &lv.back().own;
  }
  // lv may have reallocated while growing: re-point owned matrices
  for (size_t l = 1; l < lv.size(); ++l) lv[l].A = &lv[l].own;
  for (size_t l = 0; l < lv.size(); ++l) {
    AmgLevel& L = lv[l];
    L.cheb = opt.smoother == 1 && !L.isCoarsest &&
             (opt.chebLevels <= 0 || (int)l < opt.chebLevels);
    if (L.cheb) cheb_setup(L, s);
  }
  AmgLevel& Lc = lv.back();
  if (Lc.n <= 4096) dense_inverse_device(*Lc.A, Lc.coarseInvT, s);
  int maxNb = 0;
  for (size_t l = 1; l < lv.size(); ++l) {
    AmgLevel& L = lv[l];
    L.kLevel = (int)l <= opt.kcycle && !L.isCoarsest;
    if (!L.kLevel) continue;
    L.kc.alloc((size_t)L.n);
    L.kv.alloc((size_t)L.n);
    maxNb = std::max(maxNb, grid_for(L.A->rb.count, kW));
  }
  if (maxNb > 0) {
    kred.init(maxNb, s);
    kscal.alloc(8 * lv.size());
    kscal.zero(s);
  }
  AN_HIP(hipStreamSynchronize(s));
  built = true;
}

// Chebyshev data of a level from its current values: the Gershgorin bound,
// the coefficients and omdinv = (1/theta) / a_ii (and its FP32 copy)
void AmgHierarchy::cheb_setup(AmgLevel& L, hipStream_t s) {
  const DevCSR<double>& A = *L.A;
  const int nb = std::min(grid_for(L.n, kB), 1024);
  L.chebCoef.ensure(4);
  DevBuf<double> part((size_t)nb);
  k_gersh<<<nb, kB, 0, s>>>(L.n, A.rowOff.p, A.cols.p, A.vals.p, part.p);
  k_cheb_coef<<<1, kB, 0, s>>>(nb, part.p, opt.chebFraction, L.chebCoef.p);
  k_omdinv_dev<double><<<grid_for(L.n, kB), kB, 0, s>>>(L.n, A.rowOff.p, A.cols.p, A.vals.p,
                                                        L.chebCoef.p, L.omdinv.p);
  if (L.omdinv32.p)
    k_omdinv_dev<float><<<grid_for(L.n, kB), kB, 0, s>>>(L.n, A.rowOff.p, A.cols.p, A.vals.p,
                                                         L.chebCoef.p, L.omdinv32.p);
  AN_CHECK_LAUNCH();
}

bool AmgHierarchy::flexible() const {
  for (const auto& L : lv)
    if (L.kLevel) return true;
  return false;
}

void AmgHierarchy::refresh_fine(hipStream_t s) {
  if (!built) return;
  AmgLevel& L = lv[0];
  if (L.cheb) {
    PoolScope scope(&pool);
    cheb_setup(L, s);  // the bound and omdinv (and its FP32 copy) of the new values
  } else {
    k_omdinv<double><<<grid_for(L.n, kB), kB, 0, s>>>(L.n, L.A->rowOff.p, L.A->cols.p,
                                                      L.A->vals.p, opt.omega, L.omdinv.p);
    if (has32)
      k_omdinv<float><<<grid_for(L.n, kB), kB, 0, s>>>(L.n, L.A->rowOff.p, L.A->cols.p,
                                                       L.A->vals.p, opt.omega, L.omdinv32.p);
  }
  AN_CHECK_LAUNCH();
  if (has32) {
    k_cast<float><<<grid_for(L.A->nnz, kB), kB, 0, s>>>(L.A->nnz, L.A->vals.p, L.A32.vals.p);
    AN_CHECK_LAUNCH();
  }
}

void AmgHierarchy::make_fp32(hipStream_t s) {
  PoolScope scope(&pool);
  for (auto& L : lv) {
    const DevCSR<double>& A = *L.A;
    // the FP32 matrix shares the FP64 one's pattern and row blocks (A
    // outlives A32: the fine matrix outlives the hierarchy, a coarse one
    // is its own level's), so an FP32 cycle and the FP64 SpMV of precision
    // 2 stream one set of indices
    L.A32.n = A.n;
    L.A32.nnz = A.nnz;
    L.A32.rowOff.borrow(const_cast<int*>(A.rowOff.p), (size_t)A.n + 1);
    L.A32.cols.borrow(const_cast<int*>(A.cols.p), (size_t)A.nnz);
    L.A32.rb.start.borrow(const_cast<int2*>(A.rb.start.p), (size_t)A.rb.count + 1);
    L.A32.rb.count = A.rb.count;
    L.A32.has16 = A.has16;
    if (A.has16) L.A32.c16.borrow(const_cast<int16_t*>(A.c16.p), (size_t)A.nnz);
    L.A32.vals.alloc((size_t)A.nnz);
    k_cast<float><<<grid_for(A.nnz, kB), kB, 0, s>>>(A.nnz, A.vals.p, L.A32.vals.p);
    L.omdinv32.alloc((size_t)L.n);
    if (L.cheb)
      k_omdinv_dev<float><<<grid_for(L.n, kB), kB, 0, s>>>(L.n, A.rowOff.p, A.cols.p, A.vals.p,
                                                           L.chebCoef.p, L.omdinv32.p);
    else
      k_omdinv<float><<<grid_for(L.n, kB), kB, 0, s>>>(L.n, A.rowOff.p, A.cols.p, A.vals.p,
                                                       opt.omega, L.omdinv32.p);
    AN_CHECK_LAUNCH();
    L.b32.alloc((size_t)L.n);
    L.x32.alloc((size_t)L.n);
    L.r32.alloc((size_t)L.n);
    L.y32.alloc((size_t)L.n);
    if (L.kLevel) {
      L.kc32.alloc((size_t)L.n);
      L.kv32.alloc((size_t)L.n);
    }
  }
  has32 = true;
}

// x = cycle(b) on level l; y = omdinv o b (computed here when null).
// sweeps > 1: x starts at y (the first sweep from zero) and takes sweeps-1
// further Jacobi sweeps before the residual; after the coarse correction
// every sweep is a Jacobi sweep.  dots (entry level only): the last up
// kernel also finishes r.z (and z.q) of the result.
// k_level_rows on a level's matrix, with 16-bit column offsets when it has
// them
template <typename T, int KIND, int DOT = 0>
static void level_rows(const DevCSR<T>& A, int grid, hipStream_t s, const T* omd, const T* b,
                       const T* v, T* out, const T* qv, RedArgs ra, const T* wv = nullptr,
                       const double* cc = nullptr, const int* opos = nullptr) {
  if (A.has16)
    k_level_rows<T, KIND, DOT, true><<<grid, kB, 0, s>>>(A.rb.count, A.rb.start.p, A.rowOff.p,
                                                         A.cols.p, A.c16.p, A.vals.p, omd, b, v,
                                                         out, qv, ra, wv, cc, opos);
  else
    k_level_rows<T, KIND, DOT, false><<<grid, kB, 0, s>>>(A.rb.count, A.rb.start.p, A.rowOff.p,
                                                          A.cols.p, A.c16.p, A.vals.p, omd, b, v,
                                                          out, qv, ra, wv, cc, opos);
}

template <typename T>
static void kcorrect(AmgHierarchy& H, int c, hipStream_t s, int sweeps);

template <typename T>
static void vcycle(AmgHierarchy& H, int l, const T* b, const T* y, T* x, hipStream_t s,
                   int sweeps, const ApplyDots* dots = nullptr) {
  using D = LevelData<T>;
  std::vector<AmgLevel>& lv = H.lv;
  AmgLevel& L = lv[l];
  const DevCSR<T>& A = D::A(L);
  const T* qv = dots ? (const T*)dots->q : nullptr;
  if (L.isCoarsest) {
    if (L.coarseInvT.p) {
      // dense level only when n <= 4096: at most 64 chunks of either width
      const int ch = dense_chunk(L.n);
      const int nch = (L.n + ch - 1) / ch;
      const int rowBlocks = (L.n + 63) / 64;
      const int thr = 64 * std::min(nch, kDenseThreads / 64);
      if (ch == 16)
        k_dense_solve<T, 16><<<rowBlocks * spread_of(L.n), thr, 0, s>>>(L.n, L.coarseInvT.p, b, x);
      else
        k_dense_solve<T, 64><<<rowBlocks * spread_of(L.n), thr, 0, s>>>(L.n, L.coarseInvT.p, b, x);
    } else {
      T* t = D::r(L);
      fill_zero(x, sizeof(T) * (size_t)L.n, s);
      for (int it = 0; it < 20; ++it) {
        T* src = (it & 1) ? t : x;
        T* dst = (it & 1) ? x : t;
        k_jacobi<T><<<grid_for(L.n, kB), kB, 0, s>>>(L.n, A.rowOff.p, A.cols.p, A.vals.p,
                                                     D::omd(L), b, src, dst);
      }
    }
    if (dots) {
      const int nb = std::min(grid_for(L.n, kB), 1024);
      if (qv)
        k_apply_dots<T, 2><<<nb, kB, 0, s>>>(L.n, b, x, qv, dots->ra);
      else
        k_apply_dots<T, 1><<<nb, kB, 0, s>>>(L.n, b, x, qv, dots->ra);
      AN_CHECK_LAUNCH();
      finalize_parts(dots->ra, nb, qv ? 2 : 1, s);
    }
    AN_CHECK_LAUNCH();
    return;
  }
  if (!y) {
    T* yy = D::y(L);
    k_scale<T><<<grid_for(L.n, kB), kB, 0, s>>>(L.n, D::omd(L), b, yy);
    y = yy;
  }
  AmgLevel& C = lv[l + 1];
  T* r = D::r(L);
  const int G = grid_for(A.rb.count, kW);
  const int sp = spread_of(L.n);
  const RedArgs noRed{nullptr, nullptr};
  auto sweep = [&](const T* from, T* to, bool last) {  // to = from + omdinv (b - A from)
    if (last && dots && qv)
      level_rows<T, 1, 2>(A, G * sp, s, D::omd(L), b, from, to, qv, dots->ra);
    else if (last && dots)
      level_rows<T, 1, 1>(A, G * sp, s, D::omd(L), b, from, to, qv, dots->ra);
    else
      level_rows<T, 1>(A, G * sp, s, D::omd(L), b, from, to, qv, noRed);
    if (last && dots) {
      AN_CHECK_LAUNCH();
      finalize_parts(dots->ra, G, qv ? 2 : 1, s);
    }
  };
  // Chebyshev level: the cycle's y (= omdinv o b, omdinv = (1/theta) / a_ii)
  // is the first step from zero; the second goes to r, the residual of it
  // to x, e = x2 + P xc in place in r, then the two post steps from e:
  // y = e + omdinv (b - A e) and x = (cv y + cw e) + cr omdinv (b - A y)
  const double* cc = L.cheb ? L.chebCoef.p : nullptr;
  const T* xs = y;
  if (L.cheb) {
    level_rows<T, 2>(A, G * sp, s, D::omd(L), b, y, r, (const T*)nullptr, noRed,
                     (const T*)nullptr, cc);
    xs = r;
  } else {
    // pre-smoothing: x0 = y; extra sweeps ping-pong between x and r
    for (int k = 1; k < sweeps; ++k) {
      T* to = (xs == x) ? r : x;
      sweep(xs, to, false);
      xs = to;
    }
  }
  T* res = (xs == r) ? x : r;
  // Below the fine level the residual goes to res in member-list order
  // (res is read only by the restriction), which then sums contiguous runs.
  // On the fine level the scattered stores cost the down sweep more than
  // the restriction's gathers (38.9 -> 53.2 us against 12.1 -> 7.9 us at
  // 2.08 M rows), so r stays in row order there.
  const int* mpos = l > 0 ? L.mpos.p : nullptr;
  const int* mem = l > 0 ? nullptr : L.mem.p;
  level_rows<T, 0>(A, G * sp, s, D::omd(L), b, xs, res, qv, noRed, (const T*)nullptr,
                   (const double*)nullptr, mpos);
  if (L.mrb.count > 0)
    k_restrict_rb<T><<<grid_for(L.mrb.count, kW) * spread_of(L.nc), kB, 0, s>>>(
        L.mrb.count, L.mrb.start.p, L.moff.p, mem, res, D::omd(C), D::b(C), D::y(C));
  else
    k_restrict<T><<<grid_for(L.nc, kB) * spread_of(L.nc), kB, 0, s>>>(
        L.nc, L.moff.p, mem, res, D::omd(C), D::b(C), D::y(C));
  AN_CHECK_LAUNCH();
  // e = x_pre + P xc (with a K-cycle level's combination formed here),
  // then the post-smoothing sweeps ping-pong between r and x; e starts in
  // the buffer from which an odd/even number of sweeps ends in x (in place
  // when that buffer already holds x_pre)
  T* e = L.cheb ? r : (sweeps % 2 == 1) ? r : x;
  if (C.kLevel) {
    kcorrect<T>(H, l + 1, s, sweeps);
    const double* ks = H.kscal.p + 8 * (size_t)(l + 1);
    k_prolong<T, true><<<grid_for(L.n, kB) * sp, kB, 0, s>>>(L.n, xs, D::x(C), D::kc(C), ks,
                                                        L.agg.p, e);
  } else {
    vcycle<T>(H, l + 1, D::b(C), D::y(C), D::x(C), s, sweeps);
    k_prolong<T, false><<<grid_for(L.n, kB) * sp, kB, 0, s>>>(L.n, xs, D::x(C), nullptr, nullptr,
                                                         L.agg.p, e);
  }
  if (L.cheb) {
    T* x1 = D::y(L);  // y's first-step value is used up
    sweep(e, x1, false);
    if (dots && qv)
      level_rows<T, 3, 2>(A, G * sp, s, D::omd(L), b, x1, x, qv, dots->ra, e, cc);
    else if (dots)
      level_rows<T, 3, 1>(A, G * sp, s, D::omd(L), b, x1, x, qv, dots->ra, e, cc);
    else
      level_rows<T, 3>(A, G * sp, s, D::omd(L), b, x1, x, qv, noRed, e, cc);
    if (dots) {
      AN_CHECK_LAUNCH();
      finalize_parts(dots->ra, G, qv ? 2 : 1, s);
    }
    AN_CHECK_LAUNCH();
    return;
  }
  const T* cur = e;
  for (int k = 0; k < sweeps; ++k) {
    T* to = (cur == r) ? x : r;
    sweep(cur, to, k == sweeps - 1);
    cur = to;
  }
  AN_CHECK_LAUNCH();
}

// K-cycle correction on level c for the right-hand side in D::b (with y =
// omdinv o b) (Notay & Vassilevski 2008, "Recursive Krylov-based multigrid
// cycles": two flexible-CG steps preconditioned by the level's own cycle,
// always both steps):
//   c1 = B b; v1 = A c1; rho1 = c1.v1; a1 = c1.b / rho1; b -= a1 v1
//   c2 = B b; g = c2.v1; beta = c2.Ac2 - g^2/rho1; a2 = c2.b / beta
//   x = (a1 - a2 g / rho1) c1 + a2 c2
// c1 is left in D::kc, c2 in D::x and the scalars in kscal[8c..8c+5);
// the parent level's prolongation forms x.  The inner products are
// two-stage block sums like every other reduction.
template <typename T>
static void kcorrect(AmgHierarchy& H, int c, hipStream_t s, int sweeps) {
  using D = LevelData<T>;
  AmgLevel& L = H.lv[c];
  const DevCSR<T>& A = D::A(L);
  T* b = D::b(L);
  T* y = D::y(L);
  T* x = D::x(L);
  T* c1 = D::kc(L);
  T* v1 = D::kv(L);
  double* ks = H.kscal.p + 8 * (size_t)c;
  const int nb = grid_for(A.rb.count, kW);
  if (nb > H.kred.cap) throw Error(7, "K-cycle reducer too small");
  const T* none = nullptr;
  vcycle<T>(H, c, b, y, c1, s, sweeps);
  k_kprod<T, 2><<<nb, kB, 0, s>>>(A.rb.count, A.rb.start.p, A.rowOff.p, A.cols.p, A.vals.p, c1,
                                  b, none, v1, H.kred.args(ks));
  AN_CHECK_LAUNCH();
  k_kres<T><<<grid_for(L.n, kB), kB, 0, s>>>(L.n, nb, H.kred.partial.p, ks, v1, D::omd(L), b, y);
  AN_CHECK_LAUNCH();
  vcycle<T>(H, c, b, y, x, s, sweeps);
  k_kprod<T, 3><<<nb, kB, 0, s>>>(A.rb.count, A.rb.start.p, A.rowOff.p, A.cols.p, A.vals.p, x, b,
                                  v1, nullptr, H.kred.args(ks + 2));
  AN_CHECK_LAUNCH();
  finalize(H.kred, nb, 3, ks + 2, s);
}

// dots with no output (ra.out null) only carry yReady
void AmgHierarchy::apply(const double* r, double* z, hipStream_t s, const ApplyDots* dots) {
  const double* y = dots && dots->yReady ? lv[0].y.p : nullptr;
  vcycle<double>(*this, 0, r, y, z, s, std::max(opt.sweeps, 1), dots && dots->ra.out ? dots : nullptr);
}

void AmgHierarchy::apply32(const float* r, float* z, hipStream_t s, const ApplyDots* dots) {
  const float* y = dots && dots->yReady ? lv[0].y32.p : nullptr;
  vcycle<float>(*this, 0, r, y, z, s, std::max(opt.sweeps, 1), dots && dots->ra.out ? dots : nullptr);
}

}  // namespace an
