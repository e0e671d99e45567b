// simple.hip -- SIMPLE hot path of generic_simple_v1 on the GPU.
//
// Outer iteration of apps/generic_simple_v1/src/main.cu:3388-3634 with every
// field resident in HBM (the reference copies ~16 nC-sized vectors across
// PCIe per iteration).  Kernels:
//   k_mom<FULL>      momentum matrix + rhs (main.cu:1673-1896), relaxation
//                    (:1686-1695 / :1787-1794) and rAU (:1796-1802) fused,
//                    one thread per row walking its faces in face order
//   k_p_face/k_p_rows pressure matrix from rAU (:1898-1950)
//   k_rc             Rhie-Chow predicted flux (:1474-1521), one thread/face
//   k_nonorth_flux   non-orthogonal pressure flux (:1538-1593), per face
//   k_div            continuity / non-orthogonal divergence (+ pressure rhs
//                    -div* + div_no, :1595-1604), per cell in face order
//   k_correct_flux   flux correction (:1615-1658), per face
//   k_vel_p_update   p += a_p p', u = u* - rAU grad p' (:1447, 1660-1671)
//   k_force          wall-force contributions (:2828-2933) + ordered reduce
// Gathered sums run in the reference's serial face order, so a face quantity
// or a row equals the serial reference bit for bit; reductions (massRes,
// relchg, dots) use a fixed tree and are deterministic.
#include "simple.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

namespace an {

static constexpr int kB = 256;

// ----------------------------------------------------------- momentum ----

template <bool FULL>
__global__ __launch_bounds__(kB) void k_mom(MeshView m, MomArgs a, double* __restrict__ vals,
                                            double* __restrict__ rhs, double* __restrict__ rAU) {
  const int c = xcd_block(gridDim.x) * blockDim.x + threadIdx.x;
  if (c >= m.nC) return;
  const double V = m.cgeo[c].w;
  double diag = 0.0;
  double r = -a.gpc[c] * V;
  for (int k = m.cfOff[c]; k < m.cfOff[c + 1]; ++k) {
    const int e = m.cfEnt[k];
    const int f = e >> 1, side = e & 1;
    const int slotInfo = m.cfSlot[k];
    const double nd = m.fDpn[f];
    if (nd <= 1e-14) {  // main.cu:1826 -- the face contributes nothing
      if (FULL && (slotInfo & 1)) vals[slotInfo >> 1] = 0.0;
      continue;
    }
    const double af = m.fA[f];
    const double dx = m.fDx[f], dy = m.fDy[f], dz = m.fDz[f];
    const double nx = m.fNx[f], ny = m.fNy[f], nz = m.fNz[f];
    const double tcx = a.mu * (af * nx - (af / nd) * dx);
    const double tcy = a.mu * (af * ny - (af / nd) * dy);
    const double tcz = a.mu * (af * nz - (af / nd) * dz);
    const double lam = m.fLam[f];
    const int2 pn = m.fPN[f];
    const double gfx = (1.0 - lam) * a.gqx[pn.x] + lam * a.gqx[pn.y];
    const double gfy = (1.0 - lam) * a.gqy[pn.x] + lam * a.gqy[pn.y];
    const double gfz = (1.0 - lam) * a.gqz[pn.x] + lam * a.gqz[pn.y];
    const double corr = a.corrPsi * (gfx * tcx + gfy * tcy + gfz * tcz);
    if (FULL) {
      const double alpha = a.mu * af / nd;
      const double ucf = (1.0 - lam) * a.uC[pn.x] + lam * a.uC[pn.y];
      const double vcf = (1.0 - lam) * a.vC[pn.x] + lam * a.vC[pn.y];
      const double wcf = (1.0 - lam) * a.wC[pn.x] + lam * a.wC[pn.y];
      const double F = a.rho * af * (ucf * nx + vcf * ny + wcf * nz);
      const double old = (slotInfo & 1) ? 0.0 : vals[slotInfo >> 1];
      if (side == 0) {
        diag += (alpha + F * (1.0 - lam));
        vals[slotInfo >> 1] = old + (-alpha + F * lam);
      } else {
        diag += (alpha - F * lam);
        vals[slotInfo >> 1] = old + (-alpha - F * (1.0 - lam));
      }
    }
    if (side == 0)
      r += corr;
    else
      r += (-corr);
  }
  for (int j = m.cbOff[c]; j < m.cbOff[c + 1]; ++j) {
    const int b = m.cbFace[j];
    const double nd = m.bDpn[b];
    if (nd <= 1e-14) continue;
    const double af = m.bA[b];
    const double dx = m.bDx[b], dy = m.bDy[b], dz = m.bDz[b];
    const double nx = m.bNx[b], ny = m.bNy[b], nz = m.bNz[b];
    const double alpha = a.mu * af / nd;
    const double tcx = a.mu * (af * nx - (af / nd) * dx);
    const double tcy = a.mu * (af * ny - (af / nd) * dy);
    const double tcz = a.mu * (af * nz - (af / nd) * dz);
    const bool dir = a.bcT[b] == 1;
    if (dir) {
      const double g = a.qB[b];
      const double corr = a.corrPsi * (a.gqx[c] * tcx + a.gqy[c] * tcy + a.gqz[c] * tcz);
      if (FULL) diag += alpha;
      r += (alpha * g + corr);
    }
    const double ucb = dir ? a.uB[b] : a.uC[c];
    const double vcb = dir ? a.vB[b] : a.vC[c];
    const double wcb = dir ? a.wB[b] : a.wC[c];
    const double F = a.rho * af * (ucb * nx + vcb * ny + wcb * nz);
    if (dir)
      r += (-F * a.qB[b]);
    else if (FULL)
      diag += F;
  }
  const int dp = m.diagPos[c];
  if (FULL) {
    double aR = diag;
    if (a.uRelax < 0.999999) {
      const double invRelax = 1.0 / a.uRelax;
      aR = diag * invRelax;
      r += (invRelax - 1.0) * diag * a.qOld[c];
    }
    vals[dp] = aR;
    rAU[c] = (fabs(aR) > 1e-30) ? V / aR : 0.0;
  } else if (a.uRelax < 0.999999) {
    const double aPold = vals[dp] * a.uRelax;
    const double invRelax = 1.0 / a.uRelax;
    r += (invRelax - 1.0) * aPold * a.qOld[c];
  }
  rhs[c] = r;
}

// Momentum assembly for the three components at once, in two passes.
// k_mom_faces: one thread per internal face computes what both of its cells
// need, reading the face once: the non-orthogonal correction of every
// active component, corr[c][f] = corrPsi (grad_f q_c . tc), and (FULL) the
// matrix contributions dP/oP (owner side) and dN/oN (neighbour side) --
// the very expressions of k_mom, so the values carry the same bits.
// k_mom_cells: one thread per cell adds them in its incidence (face) order
// and does the boundary faces and relaxation as k_mom, giving every
// component's rhs and (FULL) the matrix and rAU bit for bit as three k_mom
// launches would (the x component takes the FULL path, y and z read the
// relaxed diagonal it produced, as their rhs-only launches do).
struct Mom3Args {
  const int* bcT;
  const double* qB[3];
  const double *uB, *vB, *wB;
  const double* gq[3][3];  // [component][x, y, z] gradient of each iterate
  const double *uC, *vC, *wC;
  const double* gpc[3];
  const double* qIt[3];    // iterates used by the relaxation
  double* rhs[3];
  double rho, mu, corrPsi, uRelax;
  int active;              // bit c: component c is assembled
};

template <bool FULL>
__global__ __launch_bounds__(kB) void k_mom_faces(MeshView m, Mom3Args a,
                                                  double* __restrict__ corr,
                                                  double* __restrict__ coef) {
  const int f = xcd_block(gridDim.x) * blockDim.x + threadIdx.x;
  if (f >= m.nIF) return;
  const double nd = m.fDpn[f];
  if (nd <= 1e-14) return;  // skipped by the cells as well
  // mu (A n - (A / n.d) d) with the precomputed T (the same value)
  const double tcx = a.mu * m.fTx[f];
  const double tcy = a.mu * m.fTy[f];
  const double tcz = a.mu * m.fTz[f];
  const double lam = m.fLam[f];
  const int2 pn = m.fPN[f];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    if (!((a.active >> c) & 1)) continue;
    const double* gx = a.gq[c][0];
    const double* gy = a.gq[c][1];
    const double* gz = a.gq[c][2];
    const double gfx = (1.0 - lam) * gx[pn.x] + lam * gx[pn.y];
    const double gfy = (1.0 - lam) * gy[pn.x] + lam * gy[pn.y];
    const double gfz = (1.0 - lam) * gz[pn.x] + lam * gz[pn.y];
    corr[(size_t)c * m.nIF + f] = a.corrPsi * (gfx * tcx + gfy * tcy + gfz * tcz);
  }
  if (FULL) {
    const double af = m.fA[f];
    const double nx = m.fNx[f], ny = m.fNy[f], nz = m.fNz[f];
    const double alpha = a.mu * af / nd;
    const double ucf = (1.0 - lam) * a.uC[pn.x] + lam * a.uC[pn.y];
    const double vcf = (1.0 - lam) * a.vC[pn.x] + lam * a.vC[pn.y];
    const double wcf = (1.0 - lam) * a.wC[pn.x] + lam * a.wC[pn.y];
    const double F = a.rho * af * (ucf * nx + vcf * ny + wcf * nz);
    const size_t n = (size_t)m.nIF;
    coef[f] = (alpha + F * (1.0 - lam));
    coef[n + f] = (-alpha + F * lam);
    coef[2 * n + f] = (alpha - F * lam);
    coef[3 * n + f] = (-alpha - F * (1.0 - lam));
  }
}

// FULL: one wave per CSR row block, lane i on row i; the block's matrix
// values are built in an LDS image and stored with unit stride (the rows'
// slot writes scattered partial cache lines).  A block whose nonzeros
// exceed the slice writes global memory directly (same order, same bits).
// rhs-only: one thread per cell.
static constexpr int kMomW = kB / 64;
template <bool FULL>
__global__ __launch_bounds__(kB) void k_mom_cells(MeshView m, Mom3Args a,
                                                  const double* __restrict__ corr,
                                                  const double* __restrict__ coef,
                                                  double* __restrict__ vals,
                                                  double* __restrict__ rAU, int nrb,
                                                  const int2* __restrict__ rbs,
                                                  double* __restrict__ dinv) {
  __shared__ double sv[FULL ? kMomW : 1][FULL ? kStreamCap : 1];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  WaveRows wr{0, 0, 0, 0, false};
  int c;
  if (FULL) {
    wr = wave_rows(nrb, rbs, xcd_block(gridDim.x) * kMomW + w);
    if (wr.r1 <= wr.r0) return;  // wave-uniform; only wave-level sync below
    c = wr.r0 + lane;
  } else {
    c = xcd_block(gridDim.x) * blockDim.x + threadIdx.x;
    if (c >= m.nC) return;
  }
  const bool stage = FULL && wr.stream;
  // the slot of CSR position k: the LDS image or global memory
  auto slotp = [&](int k) -> double* { return stage ? &sv[w][k - wr.k0] : &vals[k]; };
  if (!FULL || c < wr.r1) {
  const double V = m.cgeo[c].w;
  const size_t n = (size_t)m.nIF;
  double diag = 0.0;
  double r[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) r[q] = ((a.active >> q) & 1) ? -a.gpc[q][c] * V : 0.0;
  // faces four at a time: incidence entries, then the face values, then the
  // terms in incidence order (as one face at a time)
  constexpr int U = 4;
  const int k1 = m.cfOff[c + 1];
  for (int k0 = m.cfOff[c]; k0 < k1; k0 += U) {
    int fu[U], side[U], su[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int e = m.cfEnt[min(k0 + u, k1 - 1)];
      fu[u] = e >> 1;
      side[u] = e & 1;
      su[u] = m.cfSlot[min(k0 + u, k1 - 1)];
    }
    double nd[U], cd[U], co[U], cr[U][3];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int f = fu[u];
      nd[u] = m.fDpn[f];
      if (FULL) {
        cd[u] = coef[(size_t)(2 * side[u]) * n + f];      // diagonal share
        co[u] = coef[(size_t)(2 * side[u] + 1) * n + f];  // off-diagonal entry
      }
#pragma unroll
      for (int q = 0; q < 3; ++q) cr[u][q] = ((a.active >> q) & 1) ? corr[(size_t)q * n + f] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (k0 + u >= k1) break;
      const int slotInfo = su[u];
      if (nd[u] <= 1e-14) {  // main.cu:1826 -- the face contributes nothing
        if (FULL && (slotInfo & 1)) *slotp(slotInfo >> 1) = 0.0;
        continue;
      }
      if (FULL) {
        double* t = slotp(slotInfo >> 1);
        const double old = (slotInfo & 1) ? 0.0 : *t;
        diag += cd[u];
        *t = old + co[u];
      }
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        if (!((a.active >> q) & 1)) continue;
        if (side[u] == 0)
          r[q] += cr[u][q];
        else
          r[q] += (-cr[u][q]);
      }
    }
  }
  for (int j = m.cbOff[c]; j < m.cbOff[c + 1]; ++j) {
    const int b = m.cbFace[j];
    const double nd = m.bDpn[b];
    if (nd <= 1e-14) continue;
    const double af = m.bA[b];
    const double dx = m.bDx[b], dy = m.bDy[b], dz = m.bDz[b];
    const double nx = m.bNx[b], ny = m.bNy[b], nz = m.bNz[b];
    const double alpha = a.mu * af / nd;
    const double tcx = a.mu * (af * nx - (af / nd) * dx);
    const double tcy = a.mu * (af * ny - (af / nd) * dy);
    const double tcz = a.mu * (af * nz - (af / nd) * dz);
    const bool dir = a.bcT[b] == 1;
    if (dir) {
      if (FULL) diag += alpha;
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        if (!((a.active >> q) & 1)) continue;
        const double g = a.qB[q][b];
        const double cr =
            a.corrPsi * (a.gq[q][0][c] * tcx + a.gq[q][1][c] * tcy + a.gq[q][2][c] * tcz);
        r[q] += (alpha * g + cr);
      }
    }
    const double ucb = dir ? a.uB[b] : a.uC[c];
    const double vcb = dir ? a.vB[b] : a.vC[c];
    const double wcb = dir ? a.wB[b] : a.wC[c];
    const double F = a.rho * af * (ucb * nx + vcb * ny + wcb * nz);
    if (dir) {
#pragma unroll
      for (int q = 0; q < 3; ++q)
        if ((a.active >> q) & 1) r[q] += (-F * a.qB[q][b]);
    } else if (FULL) {
      diag += F;
    }
  }
  const int dp = m.diagPos[c];
  const bool relax = a.uRelax < 0.999999;
  const double invRelax = relax ? 1.0 / a.uRelax : 1.0;
  double aPold;
  if (FULL) {
    double aR = diag;
    if (relax) {
      aR = diag * invRelax;
      if (a.active & 1) r[0] += (invRelax - 1.0) * diag * a.qIt[0][c];
    }
    *slotp(dp) = aR;
    rAU[c] = (fabs(aR) > 1e-30) ? V / aR : 0.0;
    // the Jacobi scaling of the momentum solves (the expression of their
    // own set-up from the diagonal entry)
    dinv[c] = (aR != 0.0) ? 1.0 / aR : 1.0;
    aPold = aR * a.uRelax;
  } else {
    aPold = vals[dp] * a.uRelax;
  }
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    if (!((a.active >> q) & 1)) continue;
    if (relax && !(FULL && q == 0)) r[q] += (invRelax - 1.0) * aPold * a.qIt[q][c];
    a.rhs[q][c] = r[q];
  }
  }
  if (FULL && stage) {  // every slot of the block was written by its row
    wave_sync();
    for (int k = wr.k0 + lane; k < wr.k1; k += 64) vals[k] = sv[w][k - wr.k0];
  }
}

// ----------------------------------------------------------- pressure ----
// Pressure matrix in two passes.  k_p_face (one thread per internal face,
// unit-stride face data) computes the face coefficient A rAU_f / dpn once
// -- 0 for a face the reference skips (dpn <= 1e-14), which leaves the sums
// below bit for bit unchanged: d + 0.0 = d for the diagonal (it starts at
// +0.0 and never becomes -0.0) and old + (-0.0) = old for a slot -- and
// k_correct_flux reuses it.  k_p_rows (one wave per CSR row block) stages
// the block's incidence entries and their coefficients in LDS with
// coalesced loads, lets lane i sum row i in incidence order (the
// reference's face order) into registers, then writes the rows into an
// LDS image of the block's matrix values and stores that image with
// unit-stride writes.  A block whose entries exceed the LDS slice or that
// has a row longer than the registers hold walks global memory instead
// (same order).
__global__ __launch_bounds__(kB) void k_p_face(MeshView m, const double* __restrict__ rAU,
                                               double* __restrict__ coef) {
  const int f = xcd_block(gridDim.x) * blockDim.x + threadIdx.x;
  if (f >= m.nIF) return;
  const double dpn = m.fDpn[f];
  double c = 0.0;
  if (!(dpn <= 1e-14)) {
    const int2 pn = m.fPN[f];
    const double lam = m.fLam[f];
    const double rAUf = (1.0 - lam) * rAU[pn.x] + lam * rAU[pn.y];
    c = m.fA[f] * rAUf / dpn;
  }
  coef[f] = c;
}

static constexpr int kW = kB / 64;
static constexpr int kEnt = 512;   // staged incidence entries per wave
static constexpr int kRowReg = 8;  // row values a lane holds in registers
static_assert(kStreamCap <= kEnt && kStreamCap <= 512,
              "a stream block's image fits the slice and its offsets fit 9 bits");
__global__ __launch_bounds__(kB) void k_p_rows(MeshView m, int nrb, const int2* __restrict__ rbs,
                                               const int* __restrict__ bcP,
                                               const double* __restrict__ rAU,
                                               const double* __restrict__ coef,
                                               double* __restrict__ vals) {
  // Per wave 5 KB of LDS (8 blocks, 32 waves per CU): the coefficient of
  // each incidence entry, then -- once every lane has read its entries --
  // the block's matrix image in the same slice; the entry's slot as an
  // offset from the block's first nonzero with the first-write flag.
  __shared__ double ec[kW][kEnt];
  __shared__ unsigned short es[kW][kEnt];
  const int lb = xcd_block(gridDim.x);
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const WaveRows q = wave_rows(nrb, rbs, lb * kW + w);
  if (q.r1 <= q.r0) return;  // wave-uniform; no block barrier below
  const int c = q.r0 + lane;
  const int E0 = m.cfOff[q.r0], E1 = m.cfOff[q.r1];
  int kb = 0, len = 0;
  if (c < q.r1) {
    kb = m.rowOff[c];
    len = m.rowOff[c + 1] - kb;
  }
  // the image takes the block's nonzeros (a stream block has at most
  // kStreamCap = kEnt) and every row fits the registers
  const bool staged = q.stream && (E1 - E0) <= kEnt && __all(len <= kRowReg);
  if (staged && E1 > E0) {
    constexpr int P = kEnt / 64;
    int ent[P], slt[P];
    double cv[P];
#pragma unroll
    for (int u = 0; u < P; ++u) {
      const int k = min(E0 + lane + 64 * u, E1 - 1);
      ent[u] = m.cfEnt[k];
      slt[u] = m.cfSlot[k];
    }
#pragma unroll
    for (int u = 0; u < P; ++u) cv[u] = coef[ent[u] >> 1];
#pragma unroll
    for (int u = 0; u < P; ++u) {
      const int k = E0 + lane + 64 * u;
      if (k < E1) {
        ec[w][k - E0] = cv[u];
        es[w][k - E0] = (unsigned short)((((slt[u] >> 1) - q.k0) << 1) | (slt[u] & 1));
      }
    }
  }
  wave_sync();
  double rv[kRowReg];  // staged: the row's values, position = CSR slot - kb
#pragma unroll
  for (int r = 0; r < kRowReg; ++r) rv[r] = 0.0;  // each is written before it is read
  if (c < q.r1) {
    double diag = 0.0;
    const int k1 = m.cfOff[c + 1];
    for (int k = m.cfOff[c]; k < k1; ++k) {
      if (staged) {
        const double co = ec[w][k - E0];
        const int si = es[w][k - E0];
        const int pos = (si >> 1) - (kb - q.k0);
        diag += co;
#pragma unroll
        for (int r = 0; r < kRowReg; ++r)
          if (pos == r) rv[r] = ((si & 1) ? 0.0 : rv[r]) + (-co);
      } else {
        const double co = coef[m.cfEnt[k] >> 1];
        const int si = m.cfSlot[k];
        diag += co;
        double* t = &vals[si >> 1];
        const double old = (si & 1) ? 0.0 : *t;
        *t = old + (-co);
      }
    }
    for (int j = m.cbOff[c]; j < m.cbOff[c + 1]; ++j) {
      const int b = m.cbFace[j];
      if (bcP[b] != 1) continue;
      const double dpn = m.bDpn[b];
      if (dpn <= 1e-14) continue;
      diag += m.bA[b] * rAU[c] / dpn;
    }
    const int dp = m.diagPos[c];
    if (staged) {
#pragma unroll
      for (int r = 0; r < kRowReg; ++r)
        if (dp - kb == r) rv[r] = diag;
    } else {
      vals[dp] = diag;
    }
  }
  if (!staged) return;
  wave_sync();  // every lane has read its entries: the slice takes the image
  double* img = ec[w];
  if (c < q.r1)
#pragma unroll
    for (int r = 0; r < kRowReg; ++r)
      if (r < len) img[kb - q.k0 + r] = rv[r];
  wave_sync();
  for (int k = q.k0 + lane; k < q.k1; k += 64) vals[k] = img[k - q.k0];
}

// --------------------------------------------------------- face kernels ---
struct FaceArgs {
  const int *bcU, *bcP;
  const double *uB, *vB, *wB, *pB;
};

__global__ __launch_bounds__(kB) void k_rc(MeshView m, FaceArgs a, const double* __restrict__ u,
                                           const double* __restrict__ v,
                                           const double* __restrict__ w,
                                           const double* __restrict__ p,
                                           const double* __restrict__ gx,
                                           const double* __restrict__ gy,
                                           const double* __restrict__ gz,
                                           const double* __restrict__ rAU, double rho,
                                           double* __restrict__ phiStar) {
  const int f = xcd_block(gridDim.x) * blockDim.x + threadIdx.x;
  if (f >= m.nF) return;
  double out;
  if (f < m.nIF) {
    const int2 pn = m.fPN[f];
    const int P = pn.x, N = pn.y;
    const double dx = m.fDx[f], dy = m.fDy[f], dz = m.fDz[f];
    const double dpn = m.fDpn[f];
    const double lam = m.fLam[f];
    const double Af = m.fA[f];
    const double ubar = (1.0 - lam) * u[P] + lam * u[N];
    const double vbar = (1.0 - lam) * v[P] + lam * v[N];
    const double wbar = (1.0 - lam) * w[P] + lam * w[N];
    const double gpx = (1.0 - lam) * gx[P] + lam * gx[N];
    const double gpy = (1.0 - lam) * gy[P] + lam * gy[N];
    const double gpz = (1.0 - lam) * gz[P] + lam * gz[N];
    const double rAUf = (1.0 - lam) * rAU[P] + lam * rAU[N];
    const double phiInterp = rho * Af * (ubar * m.fNx[f] + vbar * m.fNy[f] + wbar * m.fNz[f]);
    const double rc = rho * Af * rAUf / fmax(fabs(dpn), 1.0e-30) *
                      ((p[N] - p[P]) - (gpx * dx + gpy * dy + gpz * dz));
    out = phiInterp - rc;
  } else {
    const int b = f - m.nIF;
    const int P = m.bP[b];
    const bool dir = a.bcU[b] == 1;
    const double uf = dir ? a.uB[b] : u[P];
    const double vf = dir ? a.vB[b] : v[P];
    const double wf = dir ? a.wB[b] : w[P];
    out = rho * m.bA[b] * (uf * m.bNx[b] + vf * m.bNy[b] + wf * m.bNz[b]);
  }
  // phi = phi* (main.cu:3439) is not stored: until k_correct_flux rewrites
  // every face of phi, the loop reads phi* only (its continuity included)
  phiStar[f] = out;
}

// CHECK false: no internal face has |n.d| <= 1e-30 (DevMesh::nTinyDpn), so
// the test and its load are dropped (same results)
template <bool CHECK>
__device__ __forceinline__ double nonorth_face(const MeshView& m, const FaceArgs& a,
                                               const double* __restrict__ rAU,
                                               const double* __restrict__ gx,
                                               const double* __restrict__ gy,
                                               const double* __restrict__ gz, int f) {
  double flux = 0.0;
  if (f < m.nIF) {
    if (!CHECK || fabs(m.fDpn[f]) > 1.0e-30) {
      const int2 pn = m.fPN[f];
      const double lam = m.fLam[f];
      const double rAUf = (1.0 - lam) * rAU[pn.x] + lam * rAU[pn.y];
      const double gfx = (1.0 - lam) * gx[pn.x] + lam * gx[pn.y];
      const double gfy = (1.0 - lam) * gy[pn.x] + lam * gy[pn.y];
      const double gfz = (1.0 - lam) * gz[pn.x] + lam * gz[pn.y];
      // T = A n - (A / n.d) d, precomputed with the same expressions
      flux = rAUf * (gfx * m.fTx[f] + gfy * m.fTy[f] + gfz * m.fTz[f]);
    }
  } else if (f < m.nF) {
    const int b = f - m.nIF;
    if (a.bcP[b] == 1) {
      const double dpn = m.bDpn[b];
      if (fabs(dpn) > 1.0e-30) {
        const int P = m.bP[b];
        flux = rAU[P] * (gx[P] * m.bTx[b] + gy[P] * m.bTy[b] + gz[P] * m.bTz[b]);
      }
    }
  }
  return flux;
}

// One face per thread: the kernel is bound by the gathers' cache-line
// traffic, not by loads in flight (two or four faces per thread measured
// slower).
template <bool CHECK>
__global__ __launch_bounds__(kB) void k_nonorth_flux(MeshView m, FaceArgs a,
                                                     const double* __restrict__ rAU,
                                                     const double* __restrict__ gx,
                                                     const double* __restrict__ gy,
                                                     const double* __restrict__ gz,
                                                     double* __restrict__ phiNo) {
  const int f = xcd_block(gridDim.x) * kB + threadIdx.x;
  if (f < m.nF) phiNo[f] = nonorth_face<CHECK>(m, a, rAU, gx, gy, gz, f);
}

// MODE 0: div = continuity of phi (all faces).
// MODE 1: non-orthogonal divergence of phiNo (active faces only; with CHECK
//         false every internal face is active) and the pressure rhs
//         b = -divStar + divNo.
template <int MODE, bool CHECK = true>
__global__ __launch_bounds__(kB) void k_div(MeshView m, const int* __restrict__ bcP,
                                            const double* __restrict__ flux,
                                            const double* __restrict__ divStar,
                                            double* __restrict__ out) {
  const int c = xcd_block(gridDim.x) * blockDim.x + threadIdx.x;
  if (c >= m.nC) return;
  // Latency, not bandwidth, bounds this walk (it runs at full occupancy), so
  // every load that does not depend on another is issued together: the
  // cell's bounds, divStar and -- speculatively, clamped in range -- its
  // first boundary entry; then up to six incidence entries; then all
  // their fluxes.  A cell with at most six internal faces and one
  // boundary face waits on three memory round trips.  Sums stay in face
  // order.
  const int k0 = m.cfOff[c], k1 = m.cfOff[c + 1];
  const int j0 = m.cbOff[c], j1 = m.cbOff[c + 1];
  const double ds = (MODE == 1) ? divStar[c] : 0.0;
  const int bFirst = m.nBF > 0 ? m.cbFace[min(j0, m.nBF - 1)] : 0;
  double d = 0.0;
  constexpr int U = 6;  // a hex cell's faces in one chunk (8: 5% slower)
  double bfl = 0.0;
  bool bOn = false;
  for (int kb = k0; kb < k1; kb += U) {
    int eu[U];
#pragma unroll
    for (int u = 0; u < U; ++u) eu[u] = m.cfEnt[min(kb + u, k1 - 1)];
    if (kb == k0 && j1 > j0) {  // the first boundary face's flux rides along
      bfl = flux[m.nIF + bFirst];
      bOn = MODE == 0 || (bcP[bFirst] == 1 && fabs(m.bDpn[bFirst]) > 1.0e-30);
    }
    double fl[U];
    bool on[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int f = eu[u] >> 1;
      fl[u] = flux[f];
      on[u] = MODE == 0 || !CHECK || fabs(m.fDpn[f]) > 1.0e-30;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (kb + u >= k1) break;
      if (!on[u]) continue;
      if (eu[u] & 1)
        d += -fl[u];
      else
        d += fl[u];
    }
  }
  for (int j = j0; j < j1; ++j) {
    const int b = (j == j0) ? bFirst : m.cbFace[j];
    double fv;
    bool onb;
    if (j == j0 && k1 > k0) {
      fv = bfl;
      onb = bOn;
    } else {
      fv = flux[m.nIF + b];
      onb = MODE == 0 || (bcP[b] == 1 && fabs(m.bDpn[b]) > 1.0e-30);
    }
    if (!onb) continue;
    d += fv;
  }
  out[c] = (MODE == 0) ? d : (-ds + d);
}

__global__ __launch_bounds__(kB) void k_correct_flux(MeshView m, FaceArgs a,
                                                     const double* __restrict__ rAU,
                                                     const double* __restrict__ phiStar,
                                                     const double* __restrict__ pCorr,
                                                     const double* __restrict__ phiNo,
                                                     const double* __restrict__ pcoef,
                                                     double* __restrict__ phi) {
  const int f = xcd_block(gridDim.x) * blockDim.x + threadIdx.x;
  if (f >= m.nF) return;
  if (f < m.nIF) {
    const int2 pn = m.fPN[f];
    const double dpn = m.fDpn[f];
    double coeff;
    if (!(dpn <= 1e-14)) {
      coeff = pcoef[f];  // k_p_face's A rAU_f / dpn, the same expression
    } else {
      const double lam = m.fLam[f];
      const double rAUf = (1.0 - lam) * rAU[pn.x] + lam * rAU[pn.y];
      coeff = m.fA[f] * rAUf / dpn;
    }
    phi[f] = phiStar[f] - coeff * (pCorr[pn.y] - pCorr[pn.x]) - phiNo[f];
  } else {
    const int b = f - m.nIF;
    if (a.bcP[b] == 1) {
      const double dpn = m.bDpn[b];
      if (fabs(dpn) <= 1.0e-30) {
        phi[f] = phiStar[f];
        return;
      }
      const int P = m.bP[b];
      const double coeff = m.bA[b] * rAU[P] / dpn;
      phi[f] = phiStar[f] - coeff * (a.pB[b] - pCorr[P]) - phiNo[f];
    } else {
      phi[f] = phiStar[f];
    }
  }
}

__global__ void k_p_update(int n, double* __restrict__ p, const double* __restrict__ pc,
                           double pRelax) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] += pRelax * pc[i];
}

__global__ void k_vel_correct(int n, double* __restrict__ u, double* __restrict__ v,
                              double* __restrict__ w, const double* __restrict__ rAU,
                              const double* __restrict__ gx, const double* __restrict__ gy,
                              const double* __restrict__ gz) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  u[c] = u[c] - rAU[c] * gx[c];
  v[c] = v[c] - rAU[c] * gy[c];
  w[c] = w[c] - rAU[c] * gz[c];
}

// ----------------------------------------------------------- reductions --
__device__ __forceinline__ double wsum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}
__device__ __forceinline__ double wmax(double v) {
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_down(v, off, 64));
  return v;
}

// per-block partials of sum (a-b)^2, sum a^2, max |a| (b may be null)
// Each thread takes its grid-stride elements four at a time, all loads of
// the four issued together and the sums taken in element order (the bits of
// one element at a time).
template <bool HAS_B>
__global__ __launch_bounds__(kB) void k_stats(int n, const double* __restrict__ a,
                                              const double* __restrict__ b,
                                              double* __restrict__ partial) {
  __shared__ double sh[3][kB / 64];
  double s0 = 0.0, s1 = 0.0, mx = 0.0;
  const int S = gridDim.x * kB;
  for (int i0 = blockIdx.x * kB + threadIdx.x; i0 < n; i0 += 4 * S) {
    double av[4], bv[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int i = i0 + u * S;
      av[u] = i < n ? a[i] : 0.0;
      bv[u] = (HAS_B && i < n) ? b[i] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (i0 + u * S >= n) break;
      if (HAS_B) {
        const double d = av[u] - bv[u];
        s0 += d * d;
      }
      s1 += av[u] * av[u];
      mx = fmax(mx, fabs(av[u]));
    }
  }
  s0 = wsum(s0);
  s1 = wsum(s1);
  mx = wmax(mx);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) {
    sh[0][w] = s0;
    sh[1][w] = s1;
    sh[2][w] = mx;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double t0 = 0.0, t1 = 0.0, t2 = 0.0;
    for (int i = 0; i < kB / 64; ++i) {
      t0 += sh[0][i];
      t1 += sh[1][i];
      t2 = fmax(t2, sh[2][i]);
    }
    partial[3 * blockIdx.x] = t0;
    partial[3 * blockIdx.x + 1] = t1;
    partial[3 * blockIdx.x + 2] = t2;
  }
}

// out[0..2] from the block partials of k_stats, summed in block order; the
// partials are staged in LDS by the whole block first (nparts <= kStatsMax)
static constexpr int kStatsMax = 1024;
__global__ __launch_bounds__(256) void k_stats_fin(int nparts, const double* __restrict__ partial,
                                                   double* __restrict__ out) {
  __shared__ double sh[3 * kStatsMax];
  for (int i = threadIdx.x; i < 3 * nparts; i += blockDim.x) sh[i] = partial[i];
  __syncthreads();
  if (threadIdx.x != 0) return;
  double t0 = 0.0, t1 = 0.0, t2 = 0.0;
  for (int i = 0; i < nparts; ++i) {
    t0 += sh[3 * i];
    t1 += sh[3 * i + 1];
    t2 = fmax(t2, sh[3 * i + 2]);
  }
  out[0] = t0;
  out[1] = t1;
  out[2] = t2;
}

// per-face wall-force terms (main.cu:2870-2910); stat = 12 doubles per face:
// Fp[3], Fv[3], area, used, dn, Utmag, shear, valid-dn flag
__global__ void k_force_faces(MeshView m, int f0, int nf, const double* __restrict__ u,
                              const double* __restrict__ v, const double* __restrict__ w,
                              const double* __restrict__ p, double mu, int normalSign,
                              double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nf) return;
  const int b = f0 + i - m.nIF;
  double* o = out + 12 * (size_t)i;
  for (int k = 0; k < 12; ++k) o[k] = 0.0;
  const int P = m.bP[b];
  const double A = m.bA[b];
  if (A <= 1.0e-300) return;
  double n0 = (double)normalSign * m.bNx[b];
  double n1 = (double)normalSign * m.bNy[b];
  double n2 = (double)normalSign * m.bNz[b];
  const double nmag = sqrt(n0 * n0 + n1 * n1 + n2 * n2);
  if (nmag <= 1.0e-300) return;
  const double sc = 1.0 / nmag;
  n0 = sc * n0;
  n1 = sc * n1;
  n2 = sc * n2;
  const double dx0 = -m.bDx[b], dx1 = -m.bDy[b], dx2 = -m.bDz[b];  // cc_P - xf
  double dn = dx0 * n0 + dx1 * n1 + dx2 * n2;
  if (dn <= 1.0e-14) dn = fabs(dn);
  if (dn <= 1.0e-14) dn = sqrt(dx0 * dx0 + dx1 * dx1 + dx2 * dx2);
  if (dn <= 1.0e-14) return;
  const double U0 = u[P], U1 = v[P], U2 = w[P];
  const double Un = U0 * n0 + U1 * n1 + U2 * n2;
  const double Ut0 = U0 - Un * n0, Ut1 = U1 - Un * n1, Ut2 = U2 - Un * n2;
  const double Utmag = sqrt(Ut0 * Ut0 + Ut1 * Ut1 + Ut2 * Ut2);
  const double nn[3] = {n0, n1, n2};
  const double ut[3] = {Ut0, Ut1, Ut2};
  for (int d = 0; d < 3; ++d) {
    o[d] = -p[P] * nn[d] * A;
    o[3 + d] = mu * ut[d] / dn * A;
  }
  o[6] = A;
  o[7] = 1.0;
  o[8] = dn;
  o[9] = Utmag;
  o[10] = mu * Utmag / dn;
  o[11] = 1.0;
}

// ordered serial reduction of the per-face terms (the reference's serial face
// order, main.cu:2880-2910): the block stages chunks of kForceChunk faces in
// LDS with coalesced loads, thread 0 accumulates them in face order
static constexpr int kForceChunk = 256;
__global__ __launch_bounds__(256) void k_force_reduce(int nf, const double* __restrict__ t,
                                                      double* __restrict__ out) {
  __shared__ double sh[12 * kForceChunk];
  double Fp[3] = {0, 0, 0}, Fv[3] = {0, 0, 0}, F[3] = {0, 0, 0};
  double area = 0.0, cnt = 0.0, minD = 1.0e300, maxD = 0.0, maxUt = 0.0, maxSh = 0.0;
  for (int c0 = 0; c0 < nf; c0 += kForceChunk) {
    const int nc = min(kForceChunk, nf - c0);
    __syncthreads();
    for (int i = threadIdx.x; i < 12 * nc; i += blockDim.x) sh[i] = t[12 * (size_t)c0 + i];
    __syncthreads();
    if (threadIdx.x != 0) continue;
    for (int i = 0; i < nc; ++i) {
    const double* o = sh + 12 * i;
    if (o[11] == 0.0) continue;
    for (int d = 0; d < 3; ++d) {
      Fp[d] += o[d];
      Fv[d] += o[3 + d];
      F[d] += o[d] + o[3 + d];
    }
    area += o[6];
    cnt += 1.0;
    minD = fmin(minD, o[8]);
    maxD = fmax(maxD, o[8]);
    maxUt = fmax(maxUt, o[9]);
    maxSh = fmax(maxSh, o[10]);
    }
  }
  if (threadIdx.x != 0) return;
  if (minD == 1.0e300) minD = 0.0;
  for (int d = 0; d < 3; ++d) {
    out[d] = Fp[d];
    out[3 + d] = Fv[d];
    out[6 + d] = F[d];
  }
  out[9] = area;
  out[10] = cnt;
  out[11] = minD;
  out[12] = maxD;
  out[13] = maxUt;
  out[14] = maxSh;
}

// --------------------------------------------------------------- driver --
SimpleSolver::SimpleSolver(DevMesh& mesh, const an_simple_params& p_, const int* patchUType,
                           const int* patchPType, const double* uFaceBC, const double* vFaceBC,
                           const double* wFaceBC, const double* pFaceBC, int inletPatch)
    : M(mesh), par(p_) {
  if (!M.host.hasGeometry) throw Error(AN_EINVAL, "SIMPLE needs a mesh with geometry");
  M.to_device(s);
  mv = view_of(M);
  const int nC = M.nC, nF = M.nF, nIF = M.nIF, nBF = M.nBF;
  const HostMesh& hm = M.host;
  const int nP = (int)hm.patchNames.size();
  std::vector<int> tU(nBF > 0 ? nBF : 1), tP(nBF > 0 ? nBF : 1);
  std::vector<double> vU[3], vP(nBF > 0 ? nBF : 1, 0.0);
  const double* src[3] = {uFaceBC, vFaceBC, wFaceBC};
  for (int k = 0; k < 3; ++k) vU[k].assign(nBF > 0 ? nBF : 1, 0.0);
  for (int b = 0; b < nBF; ++b) {
    const int f = nIF + b;
    const int pt = hm.bPatch[f] - 1;
    if (pt < 0 || pt >= nP) throw Error(AN_EINVAL, "boundary face without patch");
    tU[b] = patchUType[pt] == 1 ? 1 : 0;
    tP[b] = patchPType[pt] == 1 ? 1 : 0;
    for (int k = 0; k < 3; ++k) vU[k][b] = src[k][f];
    vP[b] = pFaceBC[f];
  }
  bcU.upload(tU, s);
  bcP.upload(tP, s);
  for (int k = 0; k < 3; ++k) bcUv[k].upload(vU[k], s);
  bcPv.upload(vP, s);
  // initial fields (main.cu:3302-3322): arithmetic mean of the inlet faces
  double init[3] = {0.0, 0.0, 0.0};
  if (inletPatch >= 0 && inletPatch < nP) {
    int cnt = 0;
    for (int f = nIF; f < nF; ++f)
      if (hm.bPatch[f] - 1 == inletPatch) {
        for (int k = 0; k < 3; ++k) init[k] += src[k][f];
        ++cnt;
      }
    if (cnt > 0)
      for (int k = 0; k < 3; ++k) init[k] /= cnt;
  }
  for (int k = 0; k < 3; ++k) {
    std::vector<double> h(nC, init[k]);
    q[k].upload(h, s);
    qOld[k].alloc(nC);
    gp[k].alloc(nC);
    g[k].alloc(nC);
  }
  qNew.alloc(nC);
  if (par.fuseMomentumRhs) {
    for (int c = 0; c < 3; ++c) {
      for (int d = 0; d < 3; ++d) g3[c][d].alloc(nC);
      rhs3[c].alloc(nC);
      qNew3[c].alloc(nC);
    }
    momCorr.alloc(3 * (size_t)M.nIF);
    momCoef.alloc(4 * (size_t)M.nIF);
    momDinv.alloc(nC);
  }
  p.alloc(nC);
  p.zero(s);
  pOld.alloc(nC);
  pCorr.alloc(nC);
  pCorr.zero(s);
  rAU.alloc(nC);
  rAU.zero(s);
  phiStar.alloc(nF);
  phi.alloc(nF);
  phiStar.zero(s);
  phi.zero(s);
  phiNo.alloc(nF);
  divStar.alloc(nC);
  divCorr.alloc(nC);
  divCorr.zero(s);
  rhs.alloc(nC);
  Amom.n = nC;
  Amom.nnz = M.addr.nnz;
  Amom.rowOff.alloc((size_t)nC + 1);
  Amom.cols.alloc((size_t)Amom.nnz);
  Amom.vals.alloc((size_t)Amom.nnz);
  Amom.vals.zero(s);
  AN_HIP(hipMemcpyAsync(Amom.rowOff.p, M.addr.rowOff.p, sizeof(int) * ((size_t)nC + 1),
                        hipMemcpyDeviceToDevice, s));
  AN_HIP(hipMemcpyAsync(Amom.cols.p, M.addr.cols.p, sizeof(int) * (size_t)Amom.nnz,
                        hipMemcpyDeviceToDevice, s));
  Amom.finish_pattern(s);
  pvals.alloc((size_t)Amom.nnz);
  pCoef.alloc((size_t)std::max(M.nIF, 1));
  an_pcg_options po = par.pressure;
  po.maxIter = par.pMaxit;
  po.absTol = par.pTol;
  po.relTol = par.pRelTol;
  po.precision = par.precision;
  psolver.reset(new PcgSolver(nC, M.addr.rowOff.p, M.addr.cols.p, M.addr.nnz, po, true, s));
  red.init(1024, s);
  scal.alloc(3 * kStatSlots);
  for (auto& e : ev) AN_HIP(hipEventCreate(&e));
  AN_HIP(hipStreamSynchronize(s));
}

SimpleSolver::~SimpleSolver() {
  for (auto& e : ev) (void)hipEventDestroy(e);
}

double SimpleSolver::since(int a, int b) {
  float ms = 0.f;
  AN_HIP(hipEventElapsedTime(&ms, ev[a], ev[b]));
  return ms * 1e-3;
}

void SimpleSolver::grad(int bcKind, const double* f, DevBuf<double>* out) {
  BCView bc = bcKind == 0 ? BCView{bcU.p, nullptr} : BCView{bcP.p, bcPv.p};
  launch_lsq_gradient(mv, 0, bc, f, out[0].p, out[1].p, out[2].p, s);
}

double SimpleSolver::relchg(const double* a, const double* b) {
  const int G = std::min(grid_for(M.nC, kB), 340);
  if (b)
    k_stats<true><<<G, kB, 0, s>>>(M.nC, a, b, red.partial.p);
  else
    k_stats<false><<<G, kB, 0, s>>>(M.nC, a, b, red.partial.p);
  k_stats_fin<<<1, 256, 0, s>>>(G, red.partial.p, scal.p);
  AN_CHECK_LAUNCH();
  AN_HIP(hipMemcpyAsync(hs.h, scal.p, 3 * sizeof(double), hipMemcpyDeviceToHost, s));
  AN_HIP(hipStreamSynchronize(s));
  return std::sqrt(hs.h[0] / (1.0 + hs.h[1]));
}

void SimpleSolver::queue_stats(const double* a, const double* b, int k) {
  if (k < 0 || k >= kStatSlots) throw Error(7, "statistics slot out of range");
  const int G = std::min(grid_for(M.nC, kB), 340);
  if (b)
    k_stats<true><<<G, kB, 0, s>>>(M.nC, a, b, red.partial.p);
  else
    k_stats<false><<<G, kB, 0, s>>>(M.nC, a, b, red.partial.p);
  k_stats_fin<<<1, 256, 0, s>>>(G, red.partial.p, scal.p + 3 * k);
  AN_CHECK_LAUNCH();
}

void SimpleSolver::fetch_stats(int nk) {
  AN_HIP(hipMemcpyAsync(hs.h, scal.p, 3 * (size_t)nk * sizeof(double), hipMemcpyDeviceToHost, s));
  AN_HIP(hipStreamSynchronize(s));
}

double SimpleSolver::stat_relchg(int k) const {
  return std::sqrt(hs.h[3 * k] / (1.0 + hs.h[3 * k + 1]));
}
double SimpleSolver::stat_maxabs(int k) const { return hs.h[3 * k + 2]; }

// solve_scalar_component_with_nonorth (main.cu:3350-3384)
MomArgs SimpleSolver::mom_args(int comp, const DevBuf<double>* grad) const {
  MomArgs a;
  a.bcT = bcU.p;
  a.qB = bcUv[comp].p;
  a.uB = bcUv[0].p;
  a.vB = bcUv[1].p;
  a.wB = bcUv[2].p;
  a.gqx = grad[0].p;
  a.gqy = grad[1].p;
  a.gqz = grad[2].p;
  a.uC = qOld[0].p;
  a.vC = qOld[1].p;
  a.wC = qOld[2].p;
  a.gpc = gp[comp].p;
  a.qOld = q[comp].p;
  a.rho = par.rho;
  a.mu = par.mu;
  a.corrPsi = 1.0;
  a.uRelax = par.uRelax;
  return a;
}

void SimpleSolver::momentum_component(int comp, an_simple_stats* st) {
  const int nC = M.nC;
  const int nVel = std::max(par.nVelNonOrthCorr, 0) + 1;
  // qIter = qOld ; the iterate lives in q[comp]
  AN_HIP(hipMemcpyAsync(q[comp].p, qOld[comp].p, sizeof(double) * nC, hipMemcpyDeviceToDevice, s));
  for (int it = 1; it <= nVel; ++it) {
    // velocity BC values of this component (types identical for u, v, w)
    BCView bc{bcU.p, bcUv[comp].p};
    launch_lsq_gradient(mv, 0, bc, q[comp].p, g[0].p, g[1].p, g[2].p, s);
    const MomArgs a = mom_args(comp, g);
    mark(0);
    if (comp == 0 && it == 1)
      k_mom<true><<<grid_for(nC, kB), kB, 0, s>>>(mv, a, Amom.vals.p, rhs.p, rAU.p);
    else
      k_mom<false><<<grid_for(nC, kB), kB, 0, s>>>(mv, a, Amom.vals.p, rhs.p, rAU.p);
    AN_CHECK_LAUNCH();
    mark(1);
    AN_HIP(hipMemcpyAsync(qNew.p, q[comp].p, sizeof(double) * nC, hipMemcpyDeviceToDevice, s));
    an_solve_info info{};
    an_status stt = bicgstab_device(Amom, rhs.p, qNew.p, par.velMaxit, par.velRelTol, par.velTol,
                                    bw, s, &info);
    (void)stt;  // HYPRE BiCGSTAB non-convergence is not fatal in the reference
    mark(2);
    AN_HIP(hipStreamSynchronize(s));
    tAcc[0] += since(0, 1);
    tAcc[1] += since(1, 2);
    st->bicgstabIts[comp] = info.iterations;
    double rc = relchg(qNew.p, q[comp].p);
    AN_HIP(hipMemcpyAsync(q[comp].p, qNew.p, sizeof(double) * nC, hipMemcpyDeviceToDevice, s));
    if (rc < par.corrTol) break;
  }
}

// u, v and w together (main.cu:3350-3405 run component by component; every
// component's passes read only the old fields, its own iterate and the
// matrix built in the first x pass, so the passes of the three components
// can advance in lock-step).  Each component keeps its own gradient, rhs,
// BiCGSTAB scalars, stop test and corrTol exit, so its result is bit-for-bit
// the component-by-component result; the three solves share one matrix
// stream per SpMV.
void SimpleSolver::momentum_fused(an_simple_stats* st) {
  const int nC = M.nC;
  const int nVel = std::max(par.nVelNonOrthCorr, 0) + 1;
  // the iterate starts at the old velocity (qIter = qOld): iterate() has
  // just made q[c] equal to qOld[c] and nothing has written either since
  int active = 7;
  for (int it = 1; it <= nVel && active; ++it) {
    mark(0);
    if (active == 7 && mv.gsvInv) {
      Grad3Args ga;
      ga.bcType = bcU.p;
      for (int c = 0; c < 3; ++c) {
        ga.bcVal[c] = bcUv[c].p;
        ga.phi[c] = q[c].p;
        for (int d = 0; d < 3; ++d) ga.g[c][d] = g3[c][d].p;
      }
      launch_lsq_gradient3(mv, ga, s);
    } else {
      for (int c = 0; c < 3; ++c) {
        if (!((active >> c) & 1)) continue;
        BCView bc{bcU.p, bcUv[c].p};
        launch_lsq_gradient(mv, 0, bc, q[c].p, g3[c][0].p, g3[c][1].p, g3[c][2].p, s);
      }
    }
    // the three components' rhs (and in the first pass the shared matrix,
    // built as the x component's full assembly) from one face pass and one
    // cell pass
    Mom3Args a3;
    a3.bcT = bcU.p;
    a3.uB = bcUv[0].p;
    a3.vB = bcUv[1].p;
    a3.wB = bcUv[2].p;
    a3.uC = qOld[0].p;
    a3.vC = qOld[1].p;
    a3.wC = qOld[2].p;
    for (int c = 0; c < 3; ++c) {
      a3.qB[c] = bcUv[c].p;
      for (int d = 0; d < 3; ++d) a3.gq[c][d] = g3[c][d].p;
      a3.gpc[c] = gp[c].p;
      a3.qIt[c] = q[c].p;
      a3.rhs[c] = rhs3[c].p;
    }
    a3.rho = par.rho;
    a3.mu = par.mu;
    a3.corrPsi = 1.0;
    a3.uRelax = par.uRelax;
    a3.active = active;
    const int Gf = grid_for(M.nIF, kB), Gc = grid_for(nC, kB);
    if (it == 1) {  // active == 7 here: the x component builds the matrix
      k_mom_faces<true><<<Gf, kB, 0, s>>>(mv, a3, momCorr.p, momCoef.p);
      k_mom_cells<true><<<grid_for(Amom.rb.count, kMomW), kB, 0, s>>>(
          mv, a3, momCorr.p, momCoef.p, Amom.vals.p, rAU.p, Amom.rb.count, Amom.rb.start.p,
          momDinv.p);
    } else {
      k_mom_faces<false><<<Gf, kB, 0, s>>>(mv, a3, momCorr.p, momCoef.p);
      k_mom_cells<false><<<Gc, kB, 0, s>>>(mv, a3, momCorr.p, momCoef.p, Amom.vals.p, rAU.p, 0,
                                           nullptr, nullptr);
    }
    AN_CHECK_LAUNCH();
    mark(1);
    // the solves start from the iterate q and leave the new one in qNew3
    const double* b3[3] = {rhs3[0].p, rhs3[1].p, rhs3[2].p};
    double* x3[3] = {qNew3[0].p, qNew3[1].p, qNew3[2].p};
    const double* x03[3] = {q[0].p, q[1].p, q[2].p};
    an_solve_info info[3] = {};
    (void)bicgstab3_device(Amom, b3, x3, par.velMaxit, par.velRelTol, par.velTol, bw3, s, info,
                           active, nullptr, nullptr, x03, momDinv.p);
    mark(2);
    AN_HIP(hipStreamSynchronize(s));
    tAcc[0] += since(0, 1);
    tAcc[1] += since(1, 2);
    // the components' relative changes (before q takes the new iterates),
    // then one synchronisation
    for (int c = 0; c < 3; ++c)
      if ((active >> c) & 1) queue_stats(qNew3[c].p, q[c].p, c);
    for (int c = 0; c < 3; ++c)
      if ((active >> c) & 1)
        AN_HIP(hipMemcpyAsync(q[c].p, qNew3[c].p, sizeof(double) * nC, hipMemcpyDeviceToDevice, s));
    fetch_stats(3);
    const int was = active;
    for (int c = 0; c < 3; ++c) {
      if (!((was >> c) & 1)) continue;
      st->bicgstabIts[c] = info[c].iterations;
      if (stat_relchg(c) < par.corrTol) active &= ~(1 << c);
    }
  }
}

an_status SimpleSolver::iterate(int n, an_simple_stats* last, int* convergedOut) {
  const int nC = M.nC, nF = M.nF;
  an_status status = AN_OK;
  an_simple_stats st{};
  for (int k = 0; k < n; ++k) {
    auto tw0 = std::chrono::steady_clock::now();
    ++step;
    st = an_simple_stats{};
    st.step = step;
    for (int c = 0; c < 3; ++c)
      AN_HIP(hipMemcpyAsync(qOld[c].p, q[c].p, sizeof(double) * nC, hipMemcpyDeviceToDevice, s));
    AN_HIP(hipMemcpyAsync(pOld.p, p.p, sizeof(double) * nC, hipMemcpyDeviceToDevice, s));
    mark(3);
    grad(1, pOld.p, gp);  // grad(pOld) [bcP] (main.cu:3393-3397)
    mark(4);
    if (par.fuseMomentumRhs)
      momentum_fused(&st);
    else
      for (int c = 0; c < 3; ++c) momentum_component(c, &st);
    // pressure matrix + AMG (re)setup (main.cu:3407-3411)
    mark(5);
    k_p_face<<<grid_for(M.nIF, kB), kB, 0, s>>>(mv, rAU.p, pCoef.p);
    k_p_rows<<<grid_for(Amom.rb.count, kW), kB, 0, s>>>(mv, Amom.rb.count, Amom.rb.start.p, bcP.p,
                                                       rAU.p, pCoef.p, pvals.p);
    AN_CHECK_LAUNCH();
    const int every = std::max(par.pAmgRebuildEvery, 1);
    const bool setup = (step == 1) || (((step - 1) % every) == 0) || !psolver->amg.built;
    psolver->set_values_device(pvals.p, setup);
    mark(6);
    AN_HIP(hipStreamSynchronize(s));
    tAcc[2] += since(5, 6);
    // predictor flux (grad(p) equals grad(pOld) here: p == pOld)
    mark(0);
    FaceArgs fa{bcU.p, bcP.p, bcUv[0].p, bcUv[1].p, bcUv[2].p, bcPv.p};
    k_rc<<<grid_for(nF, kB), kB, 0, s>>>(mv, fa, q[0].p, q[1].p, q[2].p, p.p, gp[0].p, gp[1].p,
                                         gp[2].p, rAU.p, par.rho, phiStar.p);
    k_div<0><<<grid_for(nC, kB), kB, 0, s>>>(mv, bcP.p, phiStar.p, nullptr, divStar.p);
    AN_CHECK_LAUNCH();
    fill_zero(pCorr.p, sizeof(double) * nC, s);
    const int nPass = std::max(par.nNonOrthCorr, 0) + 1;
    int pcgTotal = 0, pcgLast = 0;
    double tP = 0.0;
    const bool check = M.nTinyDpn > 0;
    auto nonorth_flux = [&]() {
      const int G = grid_for(nF, kB);
      if (check)
        k_nonorth_flux<true><<<G, kB, 0, s>>>(mv, fa, rAU.p, g[0].p, g[1].p, g[2].p, phiNo.p);
      else
        k_nonorth_flux<false><<<G, kB, 0, s>>>(mv, fa, rAU.p, g[0].p, g[1].p, g[2].p, phiNo.p);
    };
    for (int pass = 1; pass <= nPass; ++pass) {
      grad(1, pCorr.p, g);
      nonorth_flux();
      if (check)
        k_div<1, true><<<grid_for(nC, kB), kB, 0, s>>>(mv, bcP.p, phiNo.p, divStar.p, rhs.p);
      else
        k_div<1, false><<<grid_for(nC, kB), kB, 0, s>>>(mv, bcP.p, phiNo.p, divStar.p, rhs.p);
      AN_CHECK_LAUNCH();
      psolver->opt.relTol = (pass == nPass) ? 0.0 : par.pRelTol;
      an_solve_info info{};
      an_status ps = psolver->solve(rhs.p, pCorr.p, &info);
      if (ps == AN_ENOCONV)
        std::fprintf(stderr,
                     "WARNING: pressure PCG non-convergence. its=%d finalRel=%.6e. "
                     "Continuing with approximate pCorr.\n",
                     info.iterations, info.finalRelResNorm);
      tP += info.solveSeconds;
      pcgCum += info.iterations;
      ++pSolveCum;
      pcgLast = info.iterations;
      pcgTotal += info.iterations;
      if (pass == nPass) {
        grad(1, pCorr.p, g);
        nonorth_flux();
        k_correct_flux<<<grid_for(nF, kB), kB, 0, s>>>(mv, fa, rAU.p, phiStar.p, pCorr.p, phiNo.p,
                                                       pCoef.p, phi.p);
        k_div<0><<<grid_for(nC, kB), kB, 0, s>>>(mv, bcP.p, phi.p, nullptr, divCorr.p);
        AN_CHECK_LAUNCH();
      }
    }
    tAcc[3] += tP;
    queue_stats(divCorr.p, nullptr, 0);
    k_p_update<<<grid_for(nC, kB), kB, 0, s>>>(nC, p.p, pCorr.p, par.pRelax);
    // g still holds grad(pCorr) of the last pass (pCorr has not changed
    // since), which is what the velocity correction takes
    k_vel_correct<<<grid_for(nC, kB), kB, 0, s>>>(nC, q[0].p, q[1].p, q[2].p, rAU.p, g[0].p,
                                                  g[1].p, g[2].p);
    AN_CHECK_LAUNCH();
    queue_stats(q[0].p, qOld[0].p, 1);
    queue_stats(q[1].p, qOld[1].p, 2);
    queue_stats(q[2].p, qOld[2].p, 3);
    queue_stats(p.p, pOld.p, 4);
    fetch_stats(5);
    st.massRes = stat_maxabs(0);
    st.duRel = stat_relchg(1);
    st.dvRel = stat_relchg(2);
    st.dwRel = stat_relchg(3);
    st.dpRel = stat_relchg(4);
    st.pcgLast = pcgLast;
    st.pcgTotal = pcgTotal;
    st.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - tw0).count();
    tAcc[5] += st.seconds;
    if (!std::isfinite(st.massRes) || st.massRes > 1e50) {
      status = AN_EDIVERGED;
      break;
    }
    if (st.massRes < par.tolMass && std::max({st.duRel, st.dvRel, st.dwRel}) < par.tolVel) {
      converged = true;
      break;
    }
  }
  tAcc[4] = tAcc[5] - tAcc[0] - tAcc[1] - tAcc[2] - tAcc[3];
  if (last) *last = st;
  if (convergedOut) *convergedOut = converged ? 1 : 0;
  return status;
}

void SimpleSolver::get_field(const std::string& name, double* out, bool device) {
  const DevBuf<double>* b = nullptr;
  size_t n = (size_t)M.nC;
  if (name == "u") b = &q[0];
  else if (name == "v") b = &q[1];
  else if (name == "w") b = &q[2];
  else if (name == "p") b = &p;
  else if (name == "pCorr") b = &pCorr;
  else if (name == "rAU") b = &rAU;
  else if (name == "div") b = &divCorr;
  else if (name == "phi") { b = &phi; n = (size_t)M.nF; }
  else if (name == "pA") { b = &pvals; n = (size_t)Amom.nnz; }
  else throw Error(AN_EINVAL, "unknown field '" + name + "'");
  AN_HIP(hipMemcpy(out, b->p, sizeof(double) * n,
                   device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost));
}

void SimpleSolver::set_field(const std::string& name, const double* in, bool device) {
  DevBuf<double>* b = nullptr;
  if (name == "u") b = &q[0];
  else if (name == "v") b = &q[1];
  else if (name == "w") b = &q[2];
  else if (name == "p") b = &p;
  else throw Error(AN_EINVAL, "field '" + name + "' cannot be set");
  AN_HIP(hipMemcpy(b->p, in, sizeof(double) * M.nC,
                   device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice));
}

void SimpleSolver::patch_forces(int patch, int normalSign, double* out) {
  const HostMesh& hm = M.host;
  if (patch < 0 || patch >= (int)hm.patchNames.size()) {
    for (int i = 0; i < 15; ++i) out[i] = 0.0;
    throw Error(AN_EINVAL, "patch index out of range");
  }
  const int f0 = hm.patchStart[patch], nf = hm.patchNFaces[patch];
  DevBuf<double> t(12 * (size_t)(nf > 0 ? nf : 1)), o(15);
  if (nf > 0) {
    k_force_faces<<<grid_for(nf, kB), kB, 0, s>>>(mv, f0, nf, q[0].p, q[1].p, q[2].p, p.p,
                                                  par.mu, normalSign, t.p);
    AN_CHECK_LAUNCH();
  }
  k_force_reduce<<<1, 256, 0, s>>>(nf, t.p, o.p);
  AN_CHECK_LAUNCH();
  AN_HIP(hipMemcpyAsync(out, o.p, 15 * sizeof(double), hipMemcpyDeviceToHost, s));
  AN_HIP(hipStreamSynchronize(s));
}

void SimpleSolver::timings(double* out) {
  for (int i = 0; i < 6; ++i) out[i] = tAcc[i];
  out[6] = (double)pcgCum;
  out[7] = (double)pSolveCum;
}

}  // namespace an
