// mesh_device.hip -- uploads a HostMesh into the HBM layout described in
// an_internal.hpp and builds the GPU addressing.
#include "an_internal.hpp"
#include "kernels.hpp"

namespace an {

void DevMesh::to_device(hipStream_t s) {
  if (onDevice) return;
  const HostMesh& m = host;
  nC = m.nCells;
  nF = m.nFaces;
  nIF = m.nInternalFaces;
  nBF = nF - nIF;
  nPatches = (int)m.patchNames.size();
  addr.owner.upload(m.owner, s);
  addr.neigh.alloc((size_t)(nIF > 0 ? nIF : 1));
  if (nIF > 0) addr.neigh.upload(m.neigh, s);
  build_addressing(addr, nC, nF, nIF, addr.owner.p, addr.neigh.p, s);
  if (m.hasGeometry) {
    std::vector<double4> cg(nC);
    for (int c = 0; c < nC; ++c)
      cg[c] = make_double4(m.cc[3 * (size_t)c], m.cc[3 * (size_t)c + 1], m.cc[3 * (size_t)c + 2],
                           m.vol[c]);
    cgeo.upload(cg, s);
    const size_t ni = (size_t)(nIF > 0 ? nIF : 1);
    std::vector<int2> pn(ni);
    std::vector<double> lam(ni), dpn(ni), A(ni), nx(ni), ny(ni), nz(ni), dx(ni), dy(ni), dz(ni);
    std::vector<double> tx(ni), ty(ni), tz(ni);
    for (int f = 0; f < nIF; ++f) {
      int P = m.owner[f], N = m.neigh[f];
      pn[f] = make_int2(P, N);
      const double* cp = &m.cc[3 * (size_t)P];
      const double* cn = &m.cc[3 * (size_t)N];
      const double* x = &m.xf[3 * (size_t)f];
      const double* n = &m.nf[3 * (size_t)f];
      double d0 = cn[0] - cp[0], d1 = cn[1] - cp[1], d2 = cn[2] - cp[2];
      // interpolation weight, GSV1 main.cu:1834-1835 / LP scalar_elliptic.cu:5-12
      double denom = d0 * d0 + d1 * d1 + d2 * d2;
      double l = ((x[0] - cp[0]) * d0 + (x[1] - cp[1]) * d1 + (x[2] - cp[2]) * d2) /
                 (denom > 1e-30 ? denom : 1e-30);
      l = std::fmin(1.0, std::fmax(0.0, l));
      lam[f] = l;
      dpn[f] = n[0] * d0 + n[1] * d1 + n[2] * d2;
      A[f] = m.Af[f];
      nx[f] = n[0];
      ny[f] = n[1];
      nz[f] = n[2];
      dx[f] = d0;
      dy[f] = d1;
      dz[f] = d2;
      // T = A n - (A / n.d) d as the face kernels write it (unused when
      // |n.d| <= 1e-30: those faces are skipped)
      tx[f] = A[f] * n[0] - (A[f] / dpn[f]) * d0;
      ty[f] = A[f] * n[1] - (A[f] / dpn[f]) * d1;
      tz[f] = A[f] * n[2] - (A[f] / dpn[f]) * d2;
    }
    nTinyDpn = 0;
    for (int f = 0; f < nIF; ++f) nTinyDpn += !(std::fabs(dpn[f]) > 1.0e-30);
    fPN.upload(pn, s);
    fLam.upload(lam, s);
    fDpn.upload(dpn, s);
    fA.upload(A, s);
    fNx.upload(nx, s);
    fNy.upload(ny, s);
    fNz.upload(nz, s);
    fDx.upload(dx, s);
    fDy.upload(dy, s);
    fDz.upload(dz, s);
    fTx.upload(tx, s);
    fTy.upload(ty, s);
    fTz.upload(tz, s);
    const size_t nb = (size_t)(nBF > 0 ? nBF : 1);
    std::vector<int> bp(nb), bpa(nb);
    std::vector<double> bdpn(nb), bA_(nb), bnx(nb), bny(nb), bnz(nb), bdx(nb), bdy(nb), bdz(nb);
    std::vector<double> btx(nb), bty(nb), btz(nb);
    for (int b = 0; b < nBF; ++b) {
      int f = nIF + b;
      int P = m.owner[f];
      bp[b] = P;
      bpa[b] = m.bPatch[f] - 1;
      const double* cp = &m.cc[3 * (size_t)P];
      const double* x = &m.xf[3 * (size_t)f];
      const double* n = &m.nf[3 * (size_t)f];
      double d0 = x[0] - cp[0], d1 = x[1] - cp[1], d2 = x[2] - cp[2];
      bdx[b] = d0;
      bdy[b] = d1;
      bdz[b] = d2;
      bdpn[b] = n[0] * d0 + n[1] * d1 + n[2] * d2;
      bA_[b] = m.Af[f];
      bnx[b] = n[0];
      bny[b] = n[1];
      bnz[b] = n[2];
      btx[b] = bA_[b] * n[0] - (bA_[b] / bdpn[b]) * d0;
      bty[b] = bA_[b] * n[1] - (bA_[b] / bdpn[b]) * d1;
      btz[b] = bA_[b] * n[2] - (bA_[b] / bdpn[b]) * d2;
    }
    bP.upload(bp, s);
    bPatch.upload(bpa, s);
    bDpn.upload(bdpn, s);
    bA.upload(bA_, s);
    bNx.upload(bnx, s);
    bNy.upload(bny, s);
    bNz.upload(bnz, s);
    bDx.upload(bdx, s);
    bDy.upload(bdy, s);
    bDz.upload(bdz, s);
    bTx.upload(btx, s);
    bTy.upload(bty, s);
    bTz.upload(btz, s);
    build_gsv1_inverse(view_of(*this), gsvInv, s);
    build_lsq_operator(view_of(*this), lsqOp, s);
  }
  AN_HIP(hipStreamSynchronize(s));
  onDevice = true;
}

}  // namespace an
