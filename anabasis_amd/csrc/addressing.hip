// addressing.hip -- polyMesh owner/neighbour -> CSR addressing on the GPU.
//
// Produces exactly the pattern of build_scalar_pattern
// (LP/src/scalar_elliptic.cu:16-58 == GSV1 main.cu:1315-1340): row c holds the
// sorted unique set {c} U neighbours(c); diagPos[c]; facePP/PN/NP/NN = the
// absolute nnz slots of (P,P),(P,N),(N,P),(N,N).  Integer work only, so the
// result is bit-exact whatever the order the atomics run in: every list built
// with atomics is sorted afterwards.
//
// Also builds the cell->face incidence lists the deterministic cell-centric
// assembly kernels walk (face order, as the reference's serial face loop).
#include "an_internal.hpp"

namespace an {

// ------------------------------------------------------------- scan ----
static constexpr int kScanBlock = 256;
static constexpr int kScanItems = 8;  // per thread -> 2048 per tile
static constexpr int kScanTile = kScanBlock * kScanItems;

__device__ __forceinline__ long long block_exclusive_scan_ll(long long v, long long* sh,
                                                             long long* total) {
  // Hillis-Steele over 256 entries in LDS (one pass per power of two).
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int off = 1; off < kScanBlock; off <<= 1) {
    long long x = (t >= off) ? sh[t - off] : 0;
    __syncthreads();
    sh[t] += x;
    __syncthreads();
  }
  long long incl = sh[t];
  *total = sh[kScanBlock - 1];
  __syncthreads();
  return incl - v;
}

__global__ void k_tile_sums(const int* __restrict__ in, int n, long long* __restrict__ sums) {
  __shared__ long long sh[kScanBlock];
  const long long base = (long long)blockIdx.x * kScanTile;
  long long s = 0;
  for (int k = 0; k < kScanItems; ++k) {
    long long i = base + (long long)threadIdx.x * kScanItems + k;
    if (i < n) s += in[i];
  }
  long long tot;
  block_exclusive_scan_ll(s, sh, &tot);
  if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

__global__ void k_scan_sums(long long* sums, int nt) {
  // one block scans all tile sums in chunks of 256
  __shared__ long long sh[kScanBlock];
  long long carry = 0;
  for (int b = 0; b < nt; b += kScanBlock) {
    int i = b + threadIdx.x;
    long long v = (i < nt) ? sums[i] : 0;
    long long tot;
    long long ex = block_exclusive_scan_ll(v, sh, &tot);
    if (i < nt) sums[i] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0) sums[nt] = carry;
}

__global__ void k_tile_scan(const int* __restrict__ in, int n, const long long* __restrict__ sums,
                            int* __restrict__ out) {
  __shared__ long long sh[kScanBlock];
  const long long base = (long long)blockIdx.x * kScanTile;
  int v[kScanItems];
  long long s = 0;
  for (int k = 0; k < kScanItems; ++k) {
    long long i = base + (long long)threadIdx.x * kScanItems + k;
    v[k] = (i < n) ? in[i] : 0;
    s += v[k];
  }
  long long tot;
  long long ex = block_exclusive_scan_ll(s, sh, &tot) + sums[blockIdx.x];
  for (int k = 0; k < kScanItems; ++k) {
    long long i = base + (long long)threadIdx.x * kScanItems + k;
    if (i < n) out[i] = (int)ex;
    ex += v[k];
  }
}

__global__ void k_write_total(const long long* sums, int nt, int* out, int n) {
  out[n] = (int)sums[nt];
}

long long exclusive_scan(const int* d_in, int* d_out, int n, hipStream_t s) {
  const int nt = (n + kScanTile - 1) / kScanTile;
  DevBuf<long long> sums((size_t)nt + 1);
  if (nt > 0) {
    k_tile_sums<<<nt, kScanBlock, 0, s>>>(d_in, n, sums.p);
    AN_CHECK_LAUNCH();
  }
  k_scan_sums<<<1, kScanBlock, 0, s>>>(sums.p, nt);
  AN_CHECK_LAUNCH();
  if (nt > 0) {
    k_tile_scan<<<nt, kScanBlock, 0, s>>>(d_in, n, sums.p, d_out);
    AN_CHECK_LAUNCH();
  }
  k_write_total<<<1, 1, 0, s>>>(sums.p, nt, d_out, n);
  AN_CHECK_LAUNCH();
  long long total = 0;
  AN_HIP(hipMemcpyAsync(&total, sums.p + nt, sizeof(long long), hipMemcpyDeviceToHost, s));
  AN_HIP(hipStreamSynchronize(s));
  if (total > 0x7fffffffLL) throw Error(1, "scan total exceeds int32 range");
  return total;
}

// --------------------------------------------------------- incidence ---
__global__ void k_count_internal(int nIF, const int* __restrict__ owner,
                                 const int* __restrict__ neigh, int nC, int* __restrict__ deg,
                                 int* __restrict__ bad) {
  int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nIF) return;
  int P = owner[f], N = neigh[f];
  if (P < 0 || P >= nC || N < 0 || N >= nC || P == N) {
    atomicOr(bad, 1);
    return;
  }
  atomicAdd(&deg[P], 1);
  atomicAdd(&deg[N], 1);
}

__global__ void k_count_boundary(int nF, int nIF, const int* __restrict__ owner, int nC,
                                 int* __restrict__ deg, int* __restrict__ bad) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  int f = nIF + b;
  if (f >= nF) return;
  int P = owner[f];
  if (P < 0 || P >= nC) {
    atomicOr(bad, 1);
    return;
  }
  atomicAdd(&deg[P], 1);
}

__global__ void k_fill_internal(int nIF, const int* __restrict__ owner,
                                const int* __restrict__ neigh, int* __restrict__ cursor,
                                int* __restrict__ ent) {
  int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nIF) return;
  int P = owner[f], N = neigh[f];
  int a = atomicAdd(&cursor[P], 1);
  ent[a] = f << 1;
  int b = atomicAdd(&cursor[N], 1);
  ent[b] = (f << 1) | 1;
}

__global__ void k_fill_boundary(int nF, int nIF, const int* __restrict__ owner,
                                int* __restrict__ cursor, int* __restrict__ lst) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  int f = nIF + b;
  if (f >= nF) return;
  int a = atomicAdd(&cursor[owner[f]], 1);
  lst[a] = b;
}

__device__ __forceinline__ void isort(int* v, int n) {
  for (int i = 1; i < n; ++i) {
    int x = v[i];
    int j = i - 1;
    while (j >= 0 && v[j] > x) {
      v[j + 1] = v[j];
      --j;
    }
    v[j + 1] = x;
  }
}

// Sorts each cell's incidence segment by face; writes the neighbour cell of
// every entry into nb (same positions), sorts that copy and counts the unique
// row length (diagonal included).
__global__ void k_sort_incidence(int nC, const int* __restrict__ cfOff, int* __restrict__ ent,
                                 const int* __restrict__ owner, const int* __restrict__ neigh,
                                 int* __restrict__ nb, int* __restrict__ rowLen,
                                 const int* __restrict__ cbOff, int* __restrict__ cb) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nC) return;
  int b = cfOff[c], e = cfOff[c + 1];
  isort(ent + b, e - b);
  for (int k = b; k < e; ++k) {
    int f = ent[k] >> 1;
    nb[k] = (ent[k] & 1) ? owner[f] : neigh[f];
  }
  isort(nb + b, e - b);
  int uniq = 0, hasSelf = 0;
  for (int k = b; k < e; ++k) {
    if (k == b || nb[k] != nb[k - 1]) {
      if (nb[k] == c)
        hasSelf = 1;
      else
        ++uniq;
    }
  }
  rowLen[c] = uniq + 1;
  (void)hasSelf;
  isort(cb + cbOff[c], cbOff[c + 1] - cbOff[c]);
}

__global__ void k_fill_cols(int nC, const int* __restrict__ cfOff, const int* __restrict__ nb,
                            const int* __restrict__ rowOff, int* __restrict__ cols,
                            int* __restrict__ diagPos) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nC) return;
  int o = rowOff[c];
  int k = 0;
  bool placed = false;
  for (int i = cfOff[c]; i < cfOff[c + 1]; ++i) {
    int v = nb[i];
    if (i > cfOff[c] && v == nb[i - 1]) continue;  // duplicate pair
    if (v == c) continue;                         // degenerate self face
    if (!placed && c < v) {
      cols[o + k] = c;
      diagPos[c] = o + k;
      ++k;
      placed = true;
    }
    cols[o + k++] = v;
  }
  if (!placed) {
    cols[o + k] = c;
    diagPos[c] = o + k;
  }
}

__device__ __forceinline__ int bsearch_row(const int* __restrict__ cols, int b, int e, int key) {
  int lo = b, hi = e - 1;
  while (lo <= hi) {
    int mid = (lo + hi) >> 1;
    int v = cols[mid];
    if (v == key) return mid;
    if (v < key)
      lo = mid + 1;
    else
      hi = mid - 1;
  }
  return -1;
}

__global__ void k_face_slots(int nIF, const int* __restrict__ owner, const int* __restrict__ neigh,
                             const int* __restrict__ rowOff, const int* __restrict__ cols,
                             const int* __restrict__ diagPos, int* __restrict__ PP,
                             int* __restrict__ PN, int* __restrict__ NP, int* __restrict__ NN) {
  int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nIF) return;
  int P = owner[f], N = neigh[f];
  PP[f] = diagPos[P];
  NN[f] = diagPos[N];
  PN[f] = bsearch_row(cols, rowOff[P], rowOff[P + 1], N);
  NP[f] = bsearch_row(cols, rowOff[N], rowOff[N + 1], P);
}

__global__ void k_cell_slots(int nC, const int* __restrict__ cfOff, const int* __restrict__ ent,
                             const int* __restrict__ PN, const int* __restrict__ NP,
                             int* __restrict__ slot) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nC) return;
  int b = cfOff[c], e = cfOff[c + 1];
  for (int k = b; k < e; ++k) {
    int f = ent[k] >> 1;
    int s = (ent[k] & 1) ? NP[f] : PN[f];
    int first = 1;
    for (int j = b; j < k; ++j) {
      int g = ent[j] >> 1;
      int t = (ent[j] & 1) ? NP[g] : PN[g];
      if (t == s) {
        first = 0;
        break;
      }
    }
    slot[k] = (s << 1) | first;
  }
}

void build_addressing(DevAddressing& A, int nC, int nF, int nIF, const int* d_owner,
                      const int* d_neigh, hipStream_t s) {
  if (nC <= 0 || nF < nIF || nIF < 0) throw Error(1, "invalid mesh sizes for addressing");
  if ((long long)nIF * 2 >= (1LL << 30)) throw Error(1, "mesh too large for 32-bit addressing");
  A.nC = nC;
  A.nF = nF;
  A.nIF = nIF;
  A.nBF = nF - nIF;
  const int B = 256;
  DevBuf<int> bad(1);
  bad.zero(s);
  // internal incidence
  DevBuf<int> deg((size_t)nC + 1);
  deg.zero(s);
  if (nIF > 0) {
    k_count_internal<<<grid_for(nIF, B), B, 0, s>>>(nIF, d_owner, d_neigh, nC, deg.p, bad.p);
    AN_CHECK_LAUNCH();
  }
  DevBuf<int> bdeg((size_t)nC + 1);
  bdeg.zero(s);
  if (A.nBF > 0) {
    k_count_boundary<<<grid_for(A.nBF, B), B, 0, s>>>(nF, nIF, d_owner, nC, bdeg.p, bad.p);
    AN_CHECK_LAUNCH();
  }
  int hbad = 0;
  AN_HIP(hipMemcpyAsync(&hbad, bad.p, sizeof(int), hipMemcpyDeviceToHost, s));
  AN_HIP(hipStreamSynchronize(s));
  if (hbad) throw Error(1, "owner/neighbour index out of range");
  A.cfOff.alloc((size_t)nC + 1);
  long long nInc = exclusive_scan(deg.p, A.cfOff.p, nC, s);
  A.cbOff.alloc((size_t)nC + 1);
  long long nB = exclusive_scan(bdeg.p, A.cbOff.p, nC, s);
  if (nB != A.nBF) throw Error(7, "boundary incidence count mismatch");
  A.cfEnt.alloc((size_t)nInc + 1);
  A.cfSlot.alloc((size_t)nInc + 1);
  A.cbFace.alloc((size_t)nB + 1);
  DevBuf<int> cursor((size_t)nC + 1);
  AN_HIP(hipMemcpyAsync(cursor.p, A.cfOff.p, sizeof(int) * ((size_t)nC + 1),
                        hipMemcpyDeviceToDevice, s));
  if (nIF > 0) {
    k_fill_internal<<<grid_for(nIF, B), B, 0, s>>>(nIF, d_owner, d_neigh, cursor.p, A.cfEnt.p);
    AN_CHECK_LAUNCH();
  }
  AN_HIP(hipMemcpyAsync(cursor.p, A.cbOff.p, sizeof(int) * ((size_t)nC + 1),
                        hipMemcpyDeviceToDevice, s));
  if (A.nBF > 0) {
    k_fill_boundary<<<grid_for(A.nBF, B), B, 0, s>>>(nF, nIF, d_owner, cursor.p, A.cbFace.p);
    AN_CHECK_LAUNCH();
  }
  DevBuf<int> nb((size_t)nInc + 1);
  DevBuf<int> rowLen((size_t)nC + 1);
  k_sort_incidence<<<grid_for(nC, B), B, 0, s>>>(nC, A.cfOff.p, A.cfEnt.p, d_owner, d_neigh,
                                                 nb.p, rowLen.p, A.cbOff.p, A.cbFace.p);
  AN_CHECK_LAUNCH();
  A.rowOff.alloc((size_t)nC + 1);
  A.nnz = exclusive_scan(rowLen.p, A.rowOff.p, nC, s);
  A.cols.alloc((size_t)A.nnz);
  A.diagPos.alloc((size_t)nC);
  k_fill_cols<<<grid_for(nC, B), B, 0, s>>>(nC, A.cfOff.p, nb.p, A.rowOff.p, A.cols.p,
                                            A.diagPos.p);
  AN_CHECK_LAUNCH();
  size_t nf1 = (size_t)(nIF > 0 ? nIF : 1);
  A.facePP.alloc(nf1);
  A.facePN.alloc(nf1);
  A.faceNP.alloc(nf1);
  A.faceNN.alloc(nf1);
  if (nIF > 0) {
    k_face_slots<<<grid_for(nIF, B), B, 0, s>>>(nIF, d_owner, d_neigh, A.rowOff.p, A.cols.p,
                                                A.diagPos.p, A.facePP.p, A.facePN.p, A.faceNP.p,
                                                A.faceNN.p);
    AN_CHECK_LAUNCH();
  }
  k_cell_slots<<<grid_for(nC, B), B, 0, s>>>(nC, A.cfOff.p, A.cfEnt.p, A.facePN.p, A.faceNP.p,
                                             A.cfSlot.p);
  AN_CHECK_LAUNCH();
  AN_HIP(hipStreamSynchronize(s));
}

}  // namespace an
