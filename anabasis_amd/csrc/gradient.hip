// gradient.hip -- least-squares cell gradients, one thread per cell.
//
// GSV1 semantics (mode 0): build_lsq_gradient_operator + apply
// (apps/generic_simple_v1/src/main.cu:1159-1251, 1403-1440).  The reference
// stores a per-term operator (src, face, cx, cy, cz: 32 B/term, 429 MB at
// 128^3).  Three forms here, all with the reference's term order (sorted
// neighbours, then the cell's boundary faces in face order) and arithmetic,
// so all carry the reference's bits:
//   k_lsq_gsv1      no stored state: each thread walks its CSR row twice
//                   (normal matrix from the neighbour centroids, then each
//                   term's c_t = M^-1 (w r_t) applied);
//   k_lsq_gsv1_inv  the per-cell inverse cached with the mesh (one walk);
//   k_lsq_op        the stored per-term operator in 64-cell slices, read
//                   with unit stride across a wave (the default once a mesh
//                   is on the device; 28 B per term).
//
// LP semantics (mode 1): compute_lsq_gradient (LP/src/gradient.cu:21-83):
// Neumann boundary faces use the normal constraint r = dn n, dphi = g dn.
#include "kernels.hpp"

namespace an {


__global__ __launch_bounds__(256) void k_lsq_gsv1(MeshView m, BCView bc,
                                                  const double* __restrict__ phi,
                                                  double* __restrict__ gx, double* __restrict__ gy,
                                                  double* __restrict__ gz) {
  const int P = xcd_block(gridDim.x) * blockDim.x + threadIdx.x;
  if (P >= m.nC) return;
  const double4 xP = m.cgeo[P];
  const int rb = m.rowOff[P], re = m.rowOff[P + 1];
  const int bb = m.cbOff[P], be = m.cbOff[P + 1];
  double M[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int k = rb; k < re; ++k) {
    const int N = m.cols[k];
    if (N == P) continue;
    const double4 xN = m.cgeo[N];
    const double r0 = xN.x - xP.x, r1 = xN.y - xP.y, r2 = xN.z - xP.z;
    const double rr = r0 * r0 + r1 * r1 + r2 * r2;
    const double w = 1.0 / (rr > 1e-30 ? rr : 1e-30);
    acc_M(M, r0, r1, r2, w);
  }
  for (int j = bb; j < be; ++j) {
    const int b = m.cbFace[j];
    const double r0 = m.bDx[b], r1 = m.bDy[b], r2 = m.bDz[b];
    const double rr = r0 * r0 + r1 * r1 + r2 * r2;
    const double w = 1.0 / (rr > 1e-30 ? rr : 1e-30);
    acc_M(M, r0, r1, r2, w);
  }
  double inv[9];
  inv3(M, inv);
  const double phiP = phi[P];
  double sx = 0.0, sy = 0.0, sz = 0.0;
  for (int k = rb; k < re; ++k) {
    const int N = m.cols[k];
    if (N == P) continue;
    const double4 xN = m.cgeo[N];
    const double r0 = xN.x - xP.x, r1 = xN.y - xP.y, r2 = xN.z - xP.z;
    const double rr = r0 * r0 + r1 * r1 + r2 * r2;
    const double w = 1.0 / (rr > 1e-30 ? rr : 1e-30);
    const double bx = w * r0, by = w * r1, bz = w * r2;
    const double cx = inv[0] * bx + inv[1] * by + inv[2] * bz;
    const double cy = inv[3] * bx + inv[4] * by + inv[5] * bz;
    const double cz = inv[6] * bx + inv[7] * by + inv[8] * bz;
    const double dphi = phi[N] - phiP;
    sx += cx * dphi;
    sy += cy * dphi;
    sz += cz * dphi;
  }
  for (int j = bb; j < be; ++j) {
    const int b = m.cbFace[j];
    const double r0 = m.bDx[b], r1 = m.bDy[b], r2 = m.bDz[b];
    const double rr = r0 * r0 + r1 * r1 + r2 * r2;
    const double w = 1.0 / (rr > 1e-30 ? rr : 1e-30);
    const double bx = w * r0, by = w * r1, bz = w * r2;
    const double cx = inv[0] * bx + inv[1] * by + inv[2] * bz;
    const double cy = inv[3] * bx + inv[4] * by + inv[5] * bz;
    const double cz = inv[6] * bx + inv[7] * by + inv[8] * bz;
    const double val = (bc.type[b] == 1) ? bc.val[b] : phiP;
    const double dphi = val - phiP;
    sx += cx * dphi;
    sy += cy * dphi;
    sz += cz * dphi;
  }
  gx[P] = sx;
  gy[P] = sy;
  gz[P] = sz;
}

__global__ __launch_bounds__(256) void k_lsq_lp(MeshView m, BCView bc,
                                                const double* __restrict__ phi,
                                                double* __restrict__ gx, double* __restrict__ gy,
                                                double* __restrict__ gz) {
  const int P = xcd_block(gridDim.x) * blockDim.x + threadIdx.x;
  if (P >= m.nC) return;
  const double4 xP = m.cgeo[P];
  const double phiP = phi[P];
  double M[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  double rh0 = 0.0, rh1 = 0.0, rh2 = 0.0;
  for (int k = m.rowOff[P]; k < m.rowOff[P + 1]; ++k) {
    const int N = m.cols[k];
    if (N == P) continue;
    const double4 xN = m.cgeo[N];
    const double r0 = xN.x - xP.x, r1 = xN.y - xP.y, r2 = xN.z - xP.z;
    const double dphi = phi[N] - phiP;
    const double rr = r0 * r0 + r1 * r1 + r2 * r2;
    const double w = 1.0 / (rr > 1e-30 ? rr : 1e-30);
    acc_M(M, r0, r1, r2, w);
    rh0 += w * r0 * dphi;
    rh1 += w * r1 * dphi;
    rh2 += w * r2 * dphi;
  }
  for (int j = m.cbOff[P]; j < m.cbOff[P + 1]; ++j) {
    const int b = m.cbFace[j];
    double r0, r1, r2, dphi;
    if (bc.type[b] == 1) {
      r0 = m.bDx[b];
      r1 = m.bDy[b];
      r2 = m.bDz[b];
      dphi = bc.val[b] - phiP;
    } else {
      const double n0 = m.bNx[b], n1 = m.bNy[b], n2 = m.bNz[b];
      double dn = m.bDx[b] * n0 + m.bDy[b] * n1 + m.bDz[b] * n2;
      dn = dn > 1e-30 ? dn : 1e-30;
      r0 = dn * n0;
      r1 = dn * n1;
      r2 = dn * n2;
      dphi = bc.val[b] * dn;
    }
    const double rr = r0 * r0 + r1 * r1 + r2 * r2;
    const double w = 1.0 / (rr > 1e-30 ? rr : 1e-30);
    acc_M(M, r0, r1, r2, w);
    rh0 += w * r0 * dphi;
    rh1 += w * r1 * dphi;
    rh2 += w * r2 * dphi;
  }
  double inv[9];
  double g0 = 0.0, g1 = 0.0, g2 = 0.0;
  if (inv3(M, inv)) {
    g0 = inv[0] * rh0 + inv[1] * rh1 + inv[2] * rh2;
    g1 = inv[3] * rh0 + inv[4] * rh1 + inv[5] * rh2;
    g2 = inv[6] * rh0 + inv[7] * rh1 + inv[8] * rh2;
  }
  gx[P] = g0;
  gy[P] = g1;
  gz[P] = g2;
}

// The GSV1 normal matrix is geometry only (every neighbour and every
// boundary face enters it whatever its BC), so its inverse is computed once
// per mesh (k_lsq_gsv1's expressions: the same bits) and stored per cell as
// nine SoA arrays, read coalesced by the gradient kernel below.
__global__ __launch_bounds__(256) void k_gsv1_inverse(MeshView m, double* __restrict__ inv9) {
  const int P = blockIdx.x * blockDim.x + threadIdx.x;
  if (P >= m.nC) return;
  const double4 xP = m.cgeo[P];
  double M[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int k = m.rowOff[P]; k < m.rowOff[P + 1]; ++k) {
    const int N = m.cols[k];
    if (N == P) continue;
    const double4 xN = m.cgeo[N];
    const double r0 = xN.x - xP.x, r1 = xN.y - xP.y, r2 = xN.z - xP.z;
    const double rr = r0 * r0 + r1 * r1 + r2 * r2;
    const double w = 1.0 / (rr > 1e-30 ? rr : 1e-30);
    acc_M(M, r0, r1, r2, w);
  }
  for (int j = m.cbOff[P]; j < m.cbOff[P + 1]; ++j) {
    const int b = m.cbFace[j];
    const double r0 = m.bDx[b], r1 = m.bDy[b], r2 = m.bDz[b];
    const double rr = r0 * r0 + r1 * r1 + r2 * r2;
    const double w = 1.0 / (rr > 1e-30 ? rr : 1e-30);
    acc_M(M, r0, r1, r2, w);
  }
  double inv[9];
  inv3(M, inv);
  for (int i = 0; i < 9; ++i) inv9[(size_t)i * m.nC + P] = inv[i];
}

// NF fields sharing one BC type array, with the cached inverse: a cell's
// neighbours are taken a chunk at a time (eight for one field, so a hex or
// tet row is one chunk; four for three fields, which need the registers),
// their column indices first and then all their centroid and value gathers,
// so a thread waits on two memory round trips per chunk instead of two per
// neighbour; the boundary bounds are loaded with the row bounds.  The terms
// are added in the row's order (the diagonal skipped), as in k_lsq_gsv1.
template <int NF>
struct NbChunk {
  static constexpr int value = 4;
};
template <int NF>
struct GradFields {
  const double* phi[NF];
  const double* bval[NF];
  double* g[NF][3];
};

template <int NF>
__global__ __launch_bounds__(256) void k_lsq_gsv1_inv(MeshView m, const int* __restrict__ bcType,
                                                      GradFields<NF> a) {
  const int P = xcd_block(gridDim.x) * blockDim.x + threadIdx.x;
  if (P >= m.nC) return;
  const size_t nC = (size_t)m.nC;
  double inv[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) inv[i] = m.gsvInv[i * nC + P];
  const double4 xP = m.cgeo[P];
  double phiP[NF], sx[NF], sy[NF], sz[NF];
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    phiP[f] = a.phi[f][P];
    sx[f] = sy[f] = sz[f] = 0.0;
  }
  constexpr int kNbChunk = NbChunk<NF>::value;
  const int rb = m.rowOff[P], re = m.rowOff[P + 1];
  const int bs = m.cbOff[P], be = m.cbOff[P + 1];
  for (int k0 = rb; k0 < re; k0 += kNbChunk) {
    int nb[kNbChunk];
#pragma unroll
    for (int u = 0; u < kNbChunk; ++u) nb[u] = m.cols[min(k0 + u, re - 1)];
    double px[kNbChunk], py[kNbChunk], pz[kNbChunk], ph[kNbChunk][NF];
#pragma unroll
    for (int u = 0; u < kNbChunk; ++u) {
      // the centroid only (not the volume): fewer registers per gather
      const double* xN = reinterpret_cast<const double*>(m.cgeo + nb[u]);
      px[u] = xN[0];
      py[u] = xN[1];
      pz[u] = xN[2];
#pragma unroll
      for (int f = 0; f < NF; ++f) ph[u][f] = a.phi[f][nb[u]];
    }
#pragma unroll
    for (int u = 0; u < kNbChunk; ++u) {
      if (k0 + u >= re || nb[u] == P) continue;
      const double r0 = px[u] - xP.x, r1 = py[u] - xP.y, r2 = pz[u] - xP.z;
      const double rr = r0 * r0 + r1 * r1 + r2 * r2;
      const double w = 1.0 / (rr > 1e-30 ? rr : 1e-30);
      const double bx = w * r0, by = w * r1, bz = w * r2;
      const double cx = inv[0] * bx + inv[1] * by + inv[2] * bz;
      const double cy = inv[3] * bx + inv[4] * by + inv[5] * bz;
      const double cz = inv[6] * bx + inv[7] * by + inv[8] * bz;
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        const double dphi = ph[u][f] - phiP[f];
        sx[f] += cx * dphi;
        sy[f] += cy * dphi;
        sz[f] += cz * dphi;
      }
    }
  }
  for (int j = bs; j < be; ++j) {
    const int b = m.cbFace[j];
    const double r0 = m.bDx[b], r1 = m.bDy[b], r2 = m.bDz[b];
    const double rr = r0 * r0 + r1 * r1 + r2 * r2;
    const double w = 1.0 / (rr > 1e-30 ? rr : 1e-30);
    const double bx = w * r0, by = w * r1, bz = w * r2;
    const double cx = inv[0] * bx + inv[1] * by + inv[2] * bz;
    const double cy = inv[3] * bx + inv[4] * by + inv[5] * bz;
    const double cz = inv[6] * bx + inv[7] * by + inv[8] * bz;
    const bool fixed = bcType[b] == 1;
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      const double val = fixed ? a.bval[f][b] : phiP[f];
      const double dphi = val - phiP[f];
      sx[f] += cx * dphi;
      sy[f] += cy * dphi;
      sz[f] += cz * dphi;
    }
  }
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    a.g[f][0][P] = sx[f];
    a.g[f][1][P] = sy[f];
    a.g[f][2][P] = sz[f];
  }
}

// ------------------------------------------------- sliced operator ----
// The GSV1 gradient as the reference's stored per-term operator
// (main.cu:1159-1251: per term a source and c_t = M^-1 (w r_t); the cell's
// neighbours in CSR order, then its boundary faces in face order), in the
// slice layout of LsqOperator: a wave (64 consecutive cells) reads term j of
// all its cells with one unit-stride load per array and gathers only phi.
// The fill evaluates every coefficient with k_lsq_gsv1_inv's expressions
// from the cached inverse, so the apply gives the same bits.
__global__ void k_lsqop_len(MeshView m, int* __restrict__ len) {
  const int P = blockIdx.x * blockDim.x + threadIdx.x;
  if (P >= m.nC) return;
  int n = 0;
  for (int k = m.rowOff[P]; k < m.rowOff[P + 1]; ++k) n += m.cols[k] != P;
  len[P] = n + (m.cbOff[P + 1] - m.cbOff[P]);
}

// slots of slice s = 64 x its longest row
__global__ void k_lsqop_width(int nC, int nSlices, const int* __restrict__ len,
                              int* __restrict__ slots) {
  const int sl = blockIdx.x * blockDim.x + threadIdx.x;
  if (sl >= nSlices) return;
  int w = 0;
  for (int i = 0; i < 64; ++i) {
    const int c = 64 * sl + i;
    if (c < nC) w = max(w, len[c]);
  }
  slots[sl] = 64 * w;
}

__global__ void k_lsqop_fill(MeshView m, const int* __restrict__ off, int* __restrict__ src,
                             double* __restrict__ cxo, double* __restrict__ cyo,
                             double* __restrict__ czo) {
  const int P = blockIdx.x * blockDim.x + threadIdx.x;
  if (P >= m.nC) return;
  const size_t nC = (size_t)m.nC;
  double inv[9];
  for (int i = 0; i < 9; ++i) inv[i] = m.gsvInv[i * nC + P];
  const double4 xP = m.cgeo[P];
  long long t = (long long)off[P >> 6] + (P & 63);
  for (int k = m.rowOff[P]; k < m.rowOff[P + 1]; ++k) {
    const int N = m.cols[k];
    if (N == P) continue;
    const double4 xN = m.cgeo[N];
    const double r0 = xN.x - xP.x, r1 = xN.y - xP.y, r2 = xN.z - xP.z;
    const double rr = r0 * r0 + r1 * r1 + r2 * r2;
    const double w = 1.0 / (rr > 1e-30 ? rr : 1e-30);
    const double bx = w * r0, by = w * r1, bz = w * r2;
    src[t] = N;
    cxo[t] = inv[0] * bx + inv[1] * by + inv[2] * bz;
    cyo[t] = inv[3] * bx + inv[4] * by + inv[5] * bz;
    czo[t] = inv[6] * bx + inv[7] * by + inv[8] * bz;
    t += 64;
  }
  for (int j = m.cbOff[P]; j < m.cbOff[P + 1]; ++j) {
    const int b = m.cbFace[j];
    const double r0 = m.bDx[b], r1 = m.bDy[b], r2 = m.bDz[b];
    const double rr = r0 * r0 + r1 * r1 + r2 * r2;
    const double w = 1.0 / (rr > 1e-30 ? rr : 1e-30);
    const double bx = w * r0, by = w * r1, bz = w * r2;
    src[t] = -b - 1;
    cxo[t] = inv[0] * bx + inv[1] * by + inv[2] * bz;
    cyo[t] = inv[3] * bx + inv[4] * by + inv[5] * bz;
    czo[t] = inv[6] * bx + inv[7] * by + inv[8] * bz;
    t += 64;
  }
}

void build_lsq_operator(const MeshView& m, LsqOperator& op, hipStream_t s) {
  if (!m.gsvInv) throw Error(7, "the sliced GSV1 operator needs the cached inverse");
  const int nC = m.nC;
  const int nSl = (nC + 63) / 64;
  op.len.alloc(nC > 0 ? nC : 1);
  op.off.alloc((size_t)nSl + 1);
  DevBuf<int> slots((size_t)nSl + 1);
  k_lsqop_len<<<grid_for(nC, 256), 256, 0, s>>>(m, op.len.p);
  k_lsqop_width<<<grid_for(nSl, 256), 256, 0, s>>>(nC, nSl, op.len.p, slots.p);
  AN_CHECK_LAUNCH();
  op.slots = exclusive_scan(slots.p, op.off.p, nSl, s);
  if (op.slots > 0x7fffffffLL) throw Error(7, "LSQ operator exceeds 2^31 terms");
  const size_t ns = (size_t)(op.slots > 0 ? op.slots : 1);
  op.src.alloc(ns);
  op.cx.alloc(ns);
  op.cy.alloc(ns);
  op.cz.alloc(ns);
  // padding slots (rows shorter than their slice) are never read; cleared
  // so the arrays hold defined values
  AN_HIP(hipMemsetAsync(op.src.p, 0, sizeof(int) * ns, s));
  AN_HIP(hipMemsetAsync(op.cx.p, 0, sizeof(double) * ns, s));
  AN_HIP(hipMemsetAsync(op.cy.p, 0, sizeof(double) * ns, s));
  AN_HIP(hipMemsetAsync(op.cz.p, 0, sizeof(double) * ns, s));
  k_lsqop_fill<<<grid_for(nC, 256), 256, 0, s>>>(m, op.off.p, op.src.p, op.cx.p, op.cy.p,
                                                 op.cz.p);
  AN_CHECK_LAUNCH();
}

// Apply: per cell, terms a chunk at a time -- sources and coefficients
// (unit stride across the wave), then the phi gathers (a boundary term takes
// its face value or phi_P by the face's BC type; the value is loaded with
// the type), then the products added in term order.
template <int NF>
__global__ __launch_bounds__(256) void k_lsq_op(MeshView m, const int* __restrict__ bcType,
                                                GradFields<NF> a) {
  const int P = xcd_block(gridDim.x) * blockDim.x + threadIdx.x;
  if (P >= m.nC) return;
  const int L = m.lsqLen[P];
  // the slice (P >> 6) is the wave's own (64-thread blocks of consecutive
  // cells), so its offset is one wave-uniform load; 32-bit slot indices
  // (build_lsq_operator caps the operator at 2^31 slots)
  const int base = m.lsqOff[__builtin_amdgcn_readfirstlane(P >> 6)] + (P & 63);
  double phiP[NF], sx[NF], sy[NF], sz[NF];
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    phiP[f] = a.phi[f][P];
    sx[f] = sy[f] = sz[f] = 0.0;
  }
  // six terms per chunk for one field (a hex row in one chunk; 65 VGPRs,
  // 7 waves per SIMD -- eight terms need 81 VGPRs and run 6% slower)
  constexpr int U = NF == 1 ? 6 : 4;
  for (int j0 = 0; j0 < L; j0 += U) {
    int sv[U];
    double c0[U], c1[U], c2[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int t = base + 64 * min(j0 + u, L - 1);
      sv[u] = m.lsqSrc[t];
      c0[u] = m.lsqCx[t];
      c1[u] = m.lsqCy[t];
      c2[u] = m.lsqCz[t];
    }
    double ph[U][NF];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (sv[u] >= 0) {
#pragma unroll
        for (int f = 0; f < NF; ++f) ph[u][f] = a.phi[f][sv[u]];
      } else {
        // the face value is loaded with the BC type, not after it
        const int b = -sv[u] - 1;
        const bool fixed = bcType[b] == 1;
#pragma unroll
        for (int f = 0; f < NF; ++f) {
          const double bv = a.bval[f][b];
          ph[u][f] = fixed ? bv : phiP[f];
        }
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (j0 + u >= L) break;
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        const double dphi = ph[u][f] - phiP[f];
        sx[f] += c0[u] * dphi;
        sy[f] += c1[u] * dphi;
        sz[f] += c2[u] * dphi;
      }
    }
  }
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    a.g[f][0][P] = sx[f];
    a.g[f][1][P] = sy[f];
    a.g[f][2][P] = sz[f];
  }
}

void launch_lsq_gradient3(const MeshView& m, const Grad3Args& a, hipStream_t s) {
  if (!m.gsvInv) throw Error(7, "three-field gradient needs the cached GSV1 inverse");
  GradFields<3> f;
  for (int c = 0; c < 3; ++c) {
    f.phi[c] = a.phi[c];
    f.bval[c] = a.bcVal[c];
    for (int d = 0; d < 3; ++d) f.g[c][d] = a.g[c][d];
  }
  if (m.lsqSrc)
    k_lsq_op<3><<<grid_for(m.nC, 256), 256, 0, s>>>(m, a.bcType, f);
  else
    k_lsq_gsv1_inv<3><<<grid_for(m.nC, 256), 256, 0, s>>>(m, a.bcType, f);
  AN_CHECK_LAUNCH();
}

void build_gsv1_inverse(const MeshView& m, DevBuf<double>& inv9, hipStream_t s) {
  inv9.alloc(9 * (size_t)(m.nC > 0 ? m.nC : 1));
  k_gsv1_inverse<<<grid_for(m.nC, 256), 256, 0, s>>>(m, inv9.p);
  AN_CHECK_LAUNCH();
}

void launch_lsq_gradient(const MeshView& m, int mode, BCView bc, const double* phi, double* gx,
                         double* gy, double* gz, hipStream_t s) {
  const int B = 256;
  if (mode == 0 && m.lsqSrc)
    k_lsq_op<1><<<grid_for(m.nC, B), B, 0, s>>>(m, bc.type,
                                                 GradFields<1>{{phi}, {bc.val}, {{gx, gy, gz}}});
  else if (mode == 0 && m.gsvInv)
    k_lsq_gsv1_inv<1><<<grid_for(m.nC, B), B, 0, s>>>(m, bc.type,
                                                       GradFields<1>{{phi}, {bc.val}, {{gx, gy, gz}}});
  else if (mode == 0)
    k_lsq_gsv1<<<grid_for(m.nC, B), B, 0, s>>>(m, bc, phi, gx, gy, gz);
  else
    k_lsq_lp<<<grid_for(m.nC, B), B, 0, s>>>(m, bc, phi, gx, gy, gz);
  AN_CHECK_LAUNCH();
}

__global__ void k_aos_to_soa(int n, const double* __restrict__ a, double* __restrict__ x,
                             double* __restrict__ y, double* __restrict__ z) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  x[i] = a[3 * (size_t)i];
  y[i] = a[3 * (size_t)i + 1];
  z[i] = a[3 * (size_t)i + 2];
}
__global__ void k_soa_to_aos(int n, const double* __restrict__ x, const double* __restrict__ y,
                             const double* __restrict__ z, double* __restrict__ a) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  a[3 * (size_t)i] = x[i];
  a[3 * (size_t)i + 1] = y[i];
  a[3 * (size_t)i + 2] = z[i];
}
void launch_aos_to_soa(int n, const double* aos, double* x, double* y, double* z, hipStream_t s) {
  k_aos_to_soa<<<grid_for(n, 256), 256, 0, s>>>(n, aos, x, y, z);
  AN_CHECK_LAUNCH();
}
void launch_soa_to_aos(int n, const double* x, const double* y, const double* z, double* aos,
                       hipStream_t s) {
  k_soa_to_aos<<<grid_for(n, 256), 256, 0, s>>>(n, x, y, z, aos);
  AN_CHECK_LAUNCH();
}

}  // namespace an
