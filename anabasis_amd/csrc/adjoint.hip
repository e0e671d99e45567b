// adjoint.hip -- discrete adjoint of the steady scalar transport solve
// (libscalar::solve_steady_scalar_transport, LS/src/scalar_transport_library.cu
// :475-512, assembly :250-357).
//
// At the converged non-orthogonal fixed point the forward solution phi of
//   R(phi; gamma, bc) = A(gamma) phi - b0(gamma, bc) - c(gamma, G(phi, bc)) = 0
// (A: diffusion D_f = gamma_f |S|^2 / max(d.S, 1e-30) plus central/upwind
// convection of the frozen flux; b0: sources and boundary data; c: the
// deferred non-orthogonal correction gamma_f T_f . grad_f; G: the LP
// least-squares gradient) has, for any functional J(phi),
//   dJ/dtheta = -psi^T dR/dtheta,   (dR/dphi)^T psi = dJ/dphi.
// The adjoint is solved as the transpose of the forward fixed-point
// iteration: A^T psi_{k+1} = dJ/dphi + (dc/dphi)^T psi_k, with
// (dc/dphi)^T psi = G^T w and w_c = d(psi^T c)/d grad_c.  G^T is applied
// without storing G: z_c = M_c^-T w_c (M_c, the cell's LSQ normal matrix, is
// rebuilt from the geometry), then phibar_c = sum_n w_cn (x_c - x_n).(z_n +
// z_c) - sum_Dirichlet w_b z_c.r_b, a per-cell gather in face order.  The
// sensitivities recompute the forward's face gradients from phi.
#include "solver.hpp"
#include "kernels.hpp"

namespace an {

static constexpr int kB = 256;

// gamma_f = (1 - lam_f) k_P + lam_f k_N (internal), k_P (boundary)
__global__ void k_gamma_from_cells(MeshView m, const double* __restrict__ k,
                                   double* __restrict__ gam) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= m.nF) return;
  if (f < m.nIF) {
    const int2 pn = m.fPN[f];
    const double lam = m.fLam[f];
    gam[f] = (1.0 - lam) * k[pn.x] + lam * k[pn.y];
  } else {
    gam[f] = k[m.bP[f - m.nIF]];
  }
}

// dJ/dk_c = sum over the cell's faces of d gamma_f / d k_c * dJ/dgamma_f
__global__ void k_dk_from_dgamma(MeshView m, const double* __restrict__ dgam,
                                 double* __restrict__ dk) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= m.nC) return;
  double acc = 0.0;
  for (int k = m.cfOff[c]; k < m.cfOff[c + 1]; ++k) {
    const int e = m.cfEnt[k];
    const int f = e >> 1;
    const double lam = m.fLam[f];
    acc += ((e & 1) ? lam : (1.0 - lam)) * dgam[f];
  }
  for (int j = m.cbOff[c]; j < m.cbOff[c + 1]; ++j) acc += dgam[m.nIF + m.cbFace[j]];
  dk[c] = acc;
}

// A^T from A: the pattern is structurally symmetric, (P,N) <-> (N,P)
__global__ void k_transpose_vals(MeshView m, const double* __restrict__ a,
                                 double* __restrict__ at, const int* __restrict__ PN,
                                 const int* __restrict__ NP) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m.nIF) {
    at[PN[i]] = a[NP[i]];
    at[NP[i]] = a[PN[i]];
  }
  if (i < m.nC) at[m.diagPos[i]] = a[m.diagPos[i]];
}

// w_c = d(psi^T c)/d grad_c (the non-orthogonal correction rows)
__global__ __launch_bounds__(kB) void k_adj_w(MeshView m, BCView bc,
                                              const double* __restrict__ gam,
                                              const double* __restrict__ psi,
                                              double* __restrict__ w) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= m.nC) return;
  double w0 = 0.0, w1 = 0.0, w2 = 0.0;
  for (int k = m.cfOff[c]; k < m.cfOff[c + 1]; ++k) {
    const int e = m.cfEnt[k];
    const int f = e >> 1;
    const double g = gam[f];
    const DiffFace q = diff_face(m.fA[f], m.fNx[f], m.fNy[f], m.fNz[f], m.fDx[f], m.fDy[f],
                                 m.fDz[f], g);
    const double s = q.D / (g > 1e-30 ? g : 1e-30);
    const double T0 = q.S0 - s * q.d0, T1 = q.S1 - s * q.d1, T2 = q.S2 - s * q.d2;
    const int2 pn = m.fPN[f];
    const double lam = m.fLam[f];
    const double coef = g * (psi[pn.x] - psi[pn.y]) * ((e & 1) ? lam : (1.0 - lam));
    w0 += coef * T0;
    w1 += coef * T1;
    w2 += coef * T2;
  }
  for (int j = m.cbOff[c]; j < m.cbOff[c + 1]; ++j) {
    const int b = m.cbFace[j];
    if (bc.type[b] != 1) continue;
    const int f = m.nIF + b;
    const double g = gam[f];
    const DiffFace q = diff_face(m.bA[b], m.bNx[b], m.bNy[b], m.bNz[b], m.bDx[b], m.bDy[b],
                                 m.bDz[b], g);
    const double s = q.D / (g > 1e-30 ? g : 1e-30);
    const double coef = g * psi[c];
    w0 += coef * (q.S0 - s * q.d0);
    w1 += coef * (q.S1 - s * q.d1);
    w2 += coef * (q.S2 - s * q.d2);
  }
  w[3 * c] = w0;
  w[3 * c + 1] = w1;
  w[3 * c + 2] = w2;
}

// LSQ normal matrix of cell c exactly as k_lsq_lp builds it
__device__ __forceinline__ void lsq_matrix(const MeshView& m, BCView bc, int c, double M[9]) {
  for (int i = 0; i < 9; ++i) M[i] = 0.0;
  const double4 xP = m.cgeo[c];
  for (int k = m.rowOff[c]; k < m.rowOff[c + 1]; ++k) {
    const int N = m.cols[k];
    if (N == c) continue;
    const double4 xN = m.cgeo[N];
    const double r0 = xN.x - xP.x, r1 = xN.y - xP.y, r2 = xN.z - xP.z;
    const double rr = r0 * r0 + r1 * r1 + r2 * r2;
    acc_M(M, r0, r1, r2, 1.0 / (rr > 1e-30 ? rr : 1e-30));
  }
  for (int j = m.cbOff[c]; j < m.cbOff[c + 1]; ++j) {
    const int b = m.cbFace[j];
    double r0, r1, r2;
    if (bc.type[b] == 1) {
      r0 = m.bDx[b];
      r1 = m.bDy[b];
      r2 = m.bDz[b];
    } else {
      const double n0 = m.bNx[b], n1 = m.bNy[b], n2 = m.bNz[b];
      double dn = m.bDx[b] * n0 + m.bDy[b] * n1 + m.bDz[b] * n2;
      dn = dn > 1e-30 ? dn : 1e-30;
      r0 = dn * n0;
      r1 = dn * n1;
      r2 = dn * n2;
    }
    const double rr = r0 * r0 + r1 * r1 + r2 * r2;
    acc_M(M, r0, r1, r2, 1.0 / (rr > 1e-30 ? rr : 1e-30));
  }
}

// z_c = M_c^-T w_c (zero where the forward gradient operator is zero)
__global__ __launch_bounds__(kB) void k_adj_z(MeshView m, BCView bc, const double* __restrict__ w,
                                              double* __restrict__ z) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= m.nC) return;
  double M[9], inv[9];
  lsq_matrix(m, bc, c, M);
  inv3(M, inv);
  const double w0 = w[3 * c], w1 = w[3 * c + 1], w2 = w[3 * c + 2];
  z[3 * c] = inv[0] * w0 + inv[3] * w1 + inv[6] * w2;
  z[3 * c + 1] = inv[1] * w0 + inv[4] * w1 + inv[7] * w2;
  z[3 * c + 2] = inv[2] * w0 + inv[5] * w1 + inv[8] * w2;
}

// phibar = G^T w (cells) and its boundary-value part (per boundary face)
__global__ __launch_bounds__(kB) void k_adj_phibar(MeshView m, BCView bc,
                                                   const double* __restrict__ z,
                                                   double* __restrict__ phibar,
                                                   double* __restrict__ phibarB) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= m.nC) return;
  const double4 xc = m.cgeo[c];
  const double zc0 = z[3 * c], zc1 = z[3 * c + 1], zc2 = z[3 * c + 2];
  double acc = 0.0;
  for (int k = m.rowOff[c]; k < m.rowOff[c + 1]; ++k) {
    const int n = m.cols[k];
    if (n == c) continue;
    const double4 xn = m.cgeo[n];
    const double r0 = xn.x - xc.x, r1 = xn.y - xc.y, r2 = xn.z - xc.z;  // c -> n
    const double rr = r0 * r0 + r1 * r1 + r2 * r2;
    const double wt = 1.0 / (rr > 1e-30 ? rr : 1e-30);
    // n's term for c (r = -r) and c's term for n (r = +r)
    acc += wt * (-(z[3 * n] * r0 + z[3 * n + 1] * r1 + z[3 * n + 2] * r2) -
                 (zc0 * r0 + zc1 * r1 + zc2 * r2));
  }
  for (int j = m.cbOff[c]; j < m.cbOff[c + 1]; ++j) {
    const int b = m.cbFace[j];
    double r0, r1, r2, scale = 1.0;
    const bool dir = bc.type[b] == 1;
    if (dir) {
      r0 = m.bDx[b];
      r1 = m.bDy[b];
      r2 = m.bDz[b];
    } else {
      const double n0 = m.bNx[b], n1 = m.bNy[b], n2 = m.bNz[b];
      double dn = m.bDx[b] * n0 + m.bDy[b] * n1 + m.bDz[b] * n2;
      dn = dn > 1e-30 ? dn : 1e-30;
      r0 = dn * n0;
      r1 = dn * n1;
      r2 = dn * n2;
      scale = dn;
    }
    const double rr = r0 * r0 + r1 * r1 + r2 * r2;
    const double wt = 1.0 / (rr > 1e-30 ? rr : 1e-30);
    const double zr = wt * (zc0 * r0 + zc1 * r1 + zc2 * r2);
    if (dir) acc -= zr;
    // d grad_c / d value_b: Dirichlet (value - phi_c) term; other faces enter
    // the LP gradient as value * dn (zero-gradient faces too, as the forward
    // gradient reads their value)
    if (phibarB) phibarB[b] = zr * scale;
  }
  phibar[c] = acc;
}

// out = a + b
__global__ void k_add(int n, const double* __restrict__ a, const double* __restrict__ b,
                      double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = a[i] + b[i];
}

// dJ/dgamma per face
__global__ __launch_bounds__(kB) void k_adj_dgamma(MeshView m, BCView bc,
                                                   const double* __restrict__ gam,
                                                   const double* __restrict__ phi,
                                                   const double* __restrict__ psi,
                                                   const double* __restrict__ g, int nonOrth,
                                                   double* __restrict__ dgam) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= m.nF) return;
  if (f < m.nIF) {
    const double gm = gam[f];
    const DiffFace q = diff_face(m.fA[f], m.fNx[f], m.fNy[f], m.fNz[f], m.fDx[f], m.fDy[f],
                                 m.fDz[f], gm);
    const int2 pn = m.fPN[f];
    const double SS = q.S0 * q.S0 + q.S1 * q.S1 + q.S2 * q.S2;
    const double dS = q.d0 * q.S0 + q.d1 * q.S1 + q.d2 * q.S2;
    double term = SS / (dS > 1e-30 ? dS : 1e-30) * (phi[pn.x] - phi[pn.y]);
    if (nonOrth) {
      const double s = q.D / (gm > 1e-30 ? gm : 1e-30);
      const double lam = m.fLam[f];
      double gf[3];
      for (int d = 0; d < 3; ++d) gf[d] = (1.0 - lam) * g[3 * pn.x + d] + lam * g[3 * pn.y + d];
      term -= (q.S0 - s * q.d0) * gf[0] + (q.S1 - s * q.d1) * gf[1] + (q.S2 - s * q.d2) * gf[2];
    }
    dgam[f] = -(psi[pn.x] - psi[pn.y]) * term;
    return;
  }
  const int b = f - m.nIF;
  const int P = m.bP[b];
  const int t = bc.type[b];
  if (t == 1) {
    const double gm = gam[f];
    const DiffFace q = diff_face(m.bA[b], m.bNx[b], m.bNy[b], m.bNz[b], m.bDx[b], m.bDy[b],
                                 m.bDz[b], gm);
    const double SS = q.S0 * q.S0 + q.S1 * q.S1 + q.S2 * q.S2;
    const double dS = q.d0 * q.S0 + q.d1 * q.S1 + q.d2 * q.S2;
    double term = SS / (dS > 1e-30 ? dS : 1e-30) * (phi[P] - bc.val[b]);
    if (nonOrth) {
      const double s = q.D / (gm > 1e-30 ? gm : 1e-30);
      term -= (q.S0 - s * q.d0) * g[3 * P] + (q.S1 - s * q.d1) * g[3 * P + 1] +
              (q.S2 - s * q.d2) * g[3 * P + 2];
    }
    dgam[f] = -psi[P] * term;
  } else if (t == 0) {
    dgam[f] = psi[P] * bc.val[b] * m.bA[b];  // rhs += gamma g A (LS :351)
  } else {
    dgam[f] = 0.0;
  }
}

// dJ/d(bc value) per boundary face
__global__ void k_adj_dbc(MeshView m, BCView bc, const double* __restrict__ gam,
                          const double* __restrict__ flux, const double* __restrict__ psi,
                          const double* __restrict__ phibarB, int nonOrth,
                          double* __restrict__ dbc) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= m.nBF) return;
  const int P = m.bP[b];
  const int f = m.nIF + b;
  const double extra = nonOrth ? phibarB[b] : 0.0;
  if (bc.type[b] == 1) {
    const DiffFace q = diff_face(m.bA[b], m.bNx[b], m.bNy[b], m.bNz[b], m.bDx[b], m.bDy[b],
                                 m.bDz[b], gam[f]);
    dbc[b] = psi[P] * (q.D - flux[f]) + extra;  // rhs += D phi_B - F phi_B
  } else if (bc.type[b] == 0) {
    dbc[b] = psi[P] * gam[f] * m.bA[b] + extra;
  } else {
    dbc[b] = extra;  // the assembly ignores a zero-gradient value
  }
}

void launch_gamma_from_cells(const MeshView& m, const double* k, double* gam, hipStream_t s) {
  k_gamma_from_cells<<<grid_for(m.nF, kB), kB, 0, s>>>(m, k, gam);
  AN_CHECK_LAUNCH();
}

void launch_dk_from_dgamma(const MeshView& m, const double* dgam, double* dk, hipStream_t s) {
  k_dk_from_dgamma<<<grid_for(m.nC, kB), kB, 0, s>>>(m, dgam, dk);
  AN_CHECK_LAUNCH();
}

void launch_lsq_adjoint(const MeshView& m, BCView bc, const double* w, double* z, double* phibar,
                        double* phibarB, hipStream_t s) {
  k_adj_z<<<grid_for(m.nC, kB), kB, 0, s>>>(m, bc, w, z);
  k_adj_phibar<<<grid_for(m.nC, kB), kB, 0, s>>>(m, bc, z, phibar, phibarB);
  AN_CHECK_LAUNCH();
}

an_status transport_adjoint(const DevMesh& d, const MeshView& mv, const double* flux,
                            const double* gam, const double* Su, const double* Sp, BCView bc,
                            int upwind, int nonOrth, int nOuter, int maxIter, double relTol,
                            double absTol, const double* phi, const double* dJdphi,
                            double* dgam, double* dbc, double* psiOut, hipStream_t s,
                            an_solve_info* info) {
  const int nC = d.nC;
  // A(gamma) without the correction, and its transpose
  DevCSR<double> At;
  At.n = nC;
  At.nnz = d.addr.nnz;
  At.rowOff.alloc((size_t)nC + 1);
  At.cols.alloc((size_t)At.nnz);
  At.vals.alloc((size_t)At.nnz);
  AN_HIP(hipMemcpyAsync(At.rowOff.p, d.addr.rowOff.p, sizeof(int) * ((size_t)nC + 1),
                        hipMemcpyDeviceToDevice, s));
  AN_HIP(hipMemcpyAsync(At.cols.p, d.addr.cols.p, sizeof(int) * (size_t)At.nnz,
                        hipMemcpyDeviceToDevice, s));
  At.finish_pattern(s);
  DevBuf<double> A((size_t)At.nnz), rhs0((size_t)nC), g(3 * (size_t)nC), w(3 * (size_t)nC),
      z(3 * (size_t)nC), phibar((size_t)nC), phibarB((size_t)(d.nBF > 0 ? d.nBF : 1)),
      rhs((size_t)nC), psi((size_t)nC);
  g.zero(s);
  launch_transport_assemble(mv, flux, gam, Su, Sp, bc, g.p, g.p + nC, g.p + 2 * (size_t)nC, 0,
                            upwind, A.p, rhs0.p, s);
  k_transpose_vals<<<grid_for(std::max(d.nIF, nC), kB), kB, 0, s>>>(mv, A.p, At.vals.p,
                                                                      d.addr.facePN.p,
                                                                      d.addr.faceNP.p);
  AN_CHECK_LAUNCH();
  // forward gradient of phi (for the sensitivities), SoA -> AoS
  DevBuf<double> gs(3 * (size_t)nC);
  launch_lsq_gradient(mv, 1, bc, phi, gs.p, gs.p + nC, gs.p + 2 * (size_t)nC, s);
  launch_soa_to_aos(nC, gs.p, gs.p + nC, gs.p + 2 * (size_t)nC, g.p, s);
  // transposed fixed point
  psi.zero(s);
  BicgWork bw;
  an_solve_info last{};
  an_status st = AN_OK;
  const int passes = nonOrth ? std::max(nOuter, 1) : 1;
  for (int k = 0; k < passes; ++k) {
    const double* r = dJdphi;
    if (nonOrth && k > 0) {
      k_adj_w<<<grid_for(nC, kB), kB, 0, s>>>(mv, bc, gam, psi.p, w.p);
      launch_lsq_adjoint(mv, bc, w.p, z.p, phibar.p, nullptr, s);
      k_add<<<grid_for(nC, kB), kB, 0, s>>>(nC, dJdphi, phibar.p, rhs.p);
      AN_CHECK_LAUNCH();
      r = rhs.p;
    }
    st = bicgstab_device(At, r, psi.p, maxIter, relTol, absTol, bw, s, &last);
  }
  // sensitivities (phibarB from the converged psi)
  if (nonOrth) {
    k_adj_w<<<grid_for(nC, kB), kB, 0, s>>>(mv, bc, gam, psi.p, w.p);
    launch_lsq_adjoint(mv, bc, w.p, z.p, phibar.p, phibarB.p, s);
  }
  k_adj_dgamma<<<grid_for(d.nF, kB), kB, 0, s>>>(mv, bc, gam, phi, psi.p, g.p, nonOrth, dgam);
  if (dbc && d.nBF > 0)
    k_adj_dbc<<<grid_for(d.nBF, kB), kB, 0, s>>>(mv, bc, gam, flux, psi.p, phibarB.p, nonOrth,
                                                 dbc);
  AN_CHECK_LAUNCH();
  if (psiOut)
    AN_HIP(hipMemcpyAsync(psiOut, psi.p, sizeof(double) * nC, hipMemcpyDeviceToDevice, s));
  AN_HIP(hipStreamSynchronize(s));
  if (info) *info = last;
  return st;
}

}  // namespace an
