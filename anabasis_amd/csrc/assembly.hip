// assembly.hip -- deterministic cell-centric face-loop assembly (LP / LS).
//
// The reference assembles by a serial face loop on the host
// (LP/src/scalar_elliptic.cu:60-136, LS/src/scalar_transport_library.cu:
// 250-357) or by atomicAdd face kernels on the GPU (GSV1).  Here one thread
// owns one matrix row: it walks the cell's internal faces in face order (the
// incidence list cfEnt) and then its boundary faces in face order, computing
// each face coefficient with the reference's expressions and adding it to
// its own diagonal, off-diagonal slot and rhs.  Each value therefore receives
// its contributions in exactly the order of the reference's serial loop: no
// atomics, bitwise reproducible, and equal to the serial reference.  Each
// face is evaluated twice (once per side); the second read of its geometry
// is served by L2.
#include "kernels.hpp"

namespace an {

MeshView view_of(const DevMesh& d) {
  MeshView m;
  m.nC = d.nC;
  m.nF = d.nF;
  m.nIF = d.nIF;
  m.nBF = d.nBF;
  m.rowOff = d.addr.rowOff.p;
  m.cols = d.addr.cols.p;
  m.diagPos = d.addr.diagPos.p;
  m.cfOff = d.addr.cfOff.p;
  m.cfEnt = d.addr.cfEnt.p;
  m.cfSlot = d.addr.cfSlot.p;
  m.cbOff = d.addr.cbOff.p;
  m.cbFace = d.addr.cbFace.p;
  m.cgeo = d.cgeo.p;
  m.fPN = d.fPN.p;
  m.fLam = d.fLam.p;
  m.fDpn = d.fDpn.p;
  m.fA = d.fA.p;
  m.fNx = d.fNx.p;
  m.fNy = d.fNy.p;
  m.fNz = d.fNz.p;
  m.fDx = d.fDx.p;
  m.fDy = d.fDy.p;
  m.fDz = d.fDz.p;
  m.bP = d.bP.p;
  m.bPatch = d.bPatch.p;
  m.bDpn = d.bDpn.p;
  m.bA = d.bA.p;
  m.bNx = d.bNx.p;
  m.bNy = d.bNy.p;
  m.bNz = d.bNz.p;
  m.bDx = d.bDx.p;
  m.bDy = d.bDy.p;
  m.bDz = d.bDz.p;
  m.fTx = d.fTx.p;
  m.fTy = d.fTy.p;
  m.fTz = d.fTz.p;
  m.bTx = d.bTx.p;
  m.bTy = d.bTy.p;
  m.bTz = d.bTz.p;
  m.gsvInv = d.gsvInv.p;
  m.lsqOff = d.lsqOp.off.p;
  m.lsqLen = d.lsqOp.len.p;
  m.lsqSrc = d.lsqOp.src.p;
  m.lsqCx = d.lsqOp.cx.p;
  m.lsqCy = d.lsqOp.cy.p;
  m.lsqCz = d.lsqOp.cz.p;
  return m;
}


__device__ __forceinline__ void add_slot(double* __restrict__ vals, int slotInfo, double v) {
  // slotInfo = slot<<1 | first: the first contribution to a slot starts
  // from the reference's zeroed value.
  const int slot = slotInfo >> 1;
  const double old = (slotInfo & 1) ? 0.0 : vals[slot];
  vals[slot] = old + v;
}

__global__ __launch_bounds__(256) void k_elliptic(MeshView m, const double* __restrict__ gammaFace,
                                                  const double* __restrict__ src, BCView bc,
                                                  const double* __restrict__ gx,
                                                  const double* __restrict__ gy,
                                                  const double* __restrict__ gz, int nonOrth,
                                                  int useRef, int refCell, double refValue,
                                                  double* __restrict__ vals,
                                                  double* __restrict__ rhs) {
  const int c = xcd_block(gridDim.x) * blockDim.x + threadIdx.x;
  if (c >= m.nC) return;
  if (useRef && c == refCell) {
    // scalar_elliptic.cu:126-135: row zeroed, unit diagonal
    for (int k = m.rowOff[c]; k < m.rowOff[c + 1]; ++k) vals[k] = 0.0;
    vals[m.diagPos[c]] = 1.0;
    rhs[c] = refValue;
    return;
  }
  double diag = 0.0;
  double r = src[c] * m.cgeo[c].w;
  for (int k = m.cfOff[c]; k < m.cfOff[c + 1]; ++k) {
    const int e = m.cfEnt[k];
    const int f = e >> 1;
    const int side = e & 1;
    const double g = gammaFace[f];
    const DiffFace q =
        diff_face(m.fA[f], m.fNx[f], m.fNy[f], m.fNz[f], m.fDx[f], m.fDy[f], m.fDz[f], g);
    diag += q.D;  // PP (owner) or NN (neighbour) += D
    add_slot(vals, m.cfSlot[k], -q.D);
    if (nonOrth) {
      const int2 pn = m.fPN[f];
      const double lam = m.fLam[f];
      const double g0 = (1.0 - lam) * gx[pn.x] + lam * gx[pn.y];
      const double g1 = (1.0 - lam) * gy[pn.x] + lam * gy[pn.y];
      const double g2 = (1.0 - lam) * gz[pn.x] + lam * gz[pn.y];
      const double corr = nonorth_corr(q, g, g0, g1, g2);
      if (side == 0)
        r += corr;
      else
        r -= corr;
    }
  }
  for (int j = m.cbOff[c]; j < m.cbOff[c + 1]; ++j) {
    const int b = m.cbFace[j];
    const int f = m.nIF + b;
    const double g = gammaFace[f];
    if (bc.type[b] == 1) {
      const DiffFace q =
          diff_face(m.bA[b], m.bNx[b], m.bNy[b], m.bNz[b], m.bDx[b], m.bDy[b], m.bDz[b], g);
      diag += q.D;
      r += q.D * bc.val[b];
      if (nonOrth) r += nonorth_corr(q, g, gx[c], gy[c], gz[c]);
    } else {
      r += -g * bc.val[b] * m.bA[b];
    }
  }
  vals[m.diagPos[c]] = diag;
  rhs[c] = r;
}

__global__ __launch_bounds__(256) void k_transport(
    MeshView m, const double* __restrict__ faceFlux, const double* __restrict__ gammaFace,
    const double* __restrict__ Su, const double* __restrict__ Sp, BCView bc,
    const double* __restrict__ gx, const double* __restrict__ gy, const double* __restrict__ gz,
    int nonOrth, int upwind, double* __restrict__ vals, double* __restrict__ rhs) {
  const int c = xcd_block(gridDim.x) * blockDim.x + threadIdx.x;
  if (c >= m.nC) return;
  const double V = m.cgeo[c].w;
  double r = Su[c] * V;
  double diag = 0.0;
  diag += -Sp[c] * V;
  for (int k = m.cfOff[c]; k < m.cfOff[c + 1]; ++k) {
    const int e = m.cfEnt[k];
    const int f = e >> 1;
    const int side = e & 1;
    const int slot = m.cfSlot[k];
    const double gam = gammaFace[f];
    const DiffFace q =
        diff_face(m.fA[f], m.fNx[f], m.fNy[f], m.fNz[f], m.fDx[f], m.fDy[f], m.fDz[f], gam);
    const double lam = m.fLam[f];
    const int2 pn = m.fPN[f];
    // diffusion (scalar_transport_library.cu:291-294)
    diag += q.D;
    double off = ((slot & 1) ? 0.0 : vals[slot >> 1]) - q.D;
    if (nonOrth) {
      const double g0 = (1.0 - lam) * gx[pn.x] + lam * gx[pn.y];
      const double g1 = (1.0 - lam) * gy[pn.x] + lam * gy[pn.y];
      const double g2 = (1.0 - lam) * gz[pn.x] + lam * gz[pn.y];
      const double corr = nonorth_corr(q, gam, g0, g1, g2);
      if (side == 0)
        r += corr;
      else
        r -= corr;
    }
    // convection (:307-326)
    const double F = faceFlux[f];
    if (!upwind) {
      const double aP = F * (1.0 - lam);
      const double aN = F * lam;
      if (side == 0) {
        diag += aP;
        off += aN;
      } else {
        off -= aP;
        diag -= aN;
      }
    } else {
      if (F >= 0.0) {
        if (side == 0)
          diag += F;
        else
          off -= F;
      } else {
        if (side == 0)
          off += F;
        else
          diag -= F;
      }
    }
    vals[slot >> 1] = off;
  }
  for (int j = m.cbOff[c]; j < m.cbOff[c + 1]; ++j) {
    const int b = m.cbFace[j];
    const int f = m.nIF + b;
    const double gam = gammaFace[f];
    const DiffFace q =
        diff_face(m.bA[b], m.bNx[b], m.bNy[b], m.bNz[b], m.bDx[b], m.bDy[b], m.bDz[b], gam);
    const double F = faceFlux[f];
    if (bc.type[b] == 1) {
      const double phiB = bc.val[b];
      diag += q.D;
      r += q.D * phiB;
      if (nonOrth) r += nonorth_corr(q, gam, gx[c], gy[c], gz[c]);
      r += -F * phiB;
    } else {
      const double gradn = (bc.type[b] == 2) ? 0.0 : bc.val[b];
      r += gam * gradn * m.bA[b];
      diag += F;
    }
  }
  vals[m.diagPos[c]] = diag;
  rhs[c] = r;
}

void launch_elliptic_assemble(const MeshView& m, const double* gammaFace,
                              const double* cellSource, BCView bc, const double* gx,
                              const double* gy, const double* gz, int includeNonOrth,
                              int useRef, int refCell, double refValue, double* vals,
                              double* rhs, hipStream_t s) {
  k_elliptic<<<grid_for(m.nC, 256), 256, 0, s>>>(m, gammaFace, cellSource, bc, gx, gy, gz,
                                                  includeNonOrth, useRef, refCell, refValue, vals,
                                                  rhs);
  AN_CHECK_LAUNCH();
}

void launch_transport_assemble(const MeshView& m, const double* faceFlux, const double* gammaFace,
                               const double* Su, const double* Sp, BCView bc, const double* gx,
                               const double* gy, const double* gz, int includeNonOrth,
                               int upwind, double* vals, double* rhs, hipStream_t s) {
  k_transport<<<grid_for(m.nC, 256), 256, 0, s>>>(m, faceFlux, gammaFace, Su, Sp, bc, gx, gy, gz,
                                                   includeNonOrth, upwind, vals, rhs);
  AN_CHECK_LAUNCH();
}

}  // namespace an
