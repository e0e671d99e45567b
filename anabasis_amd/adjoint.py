"""Discrete adjoint of the steady scalar transport solve (C ABI
an_transport_adjoint) and a torch.autograd.Function around it.

Forward model (libscalar::solve_steady_scalar_transport,
LS/src/scalar_transport_library.cu:475-512) with a per-cell diffusivity k
mapped to faces by an_face_gamma_from_cells; parameters are k[nCells] and
the boundary values bcValBF[nBF] (Dirichlet values / Neumann gradients).
The backward pass is one transposed solve sequence on the GPU; nothing of
the forward is stored except phi.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from ._lib import check, f64, i32, lib, ptr


class TransportProblem:
    """Fixed data of one transport problem: mesh, frozen face flux, sources,
    boundary types and solver options."""

    def __init__(self, mesh, face_flux, bc_types, Su=None, Sp=None, upwind=False, nonOrth=True,
                 nNonOrthCorr=20, maxIter=4000, relTol=0.0, absTol=1e-13):
        self.mesh = mesh
        self.flux = f64(face_flux)
        self.bc_types = i32(bc_types)
        self.Su = f64(np.zeros(mesh.nCells) if Su is None else Su)
        self.Sp = f64(np.zeros(mesh.nCells) if Sp is None else Sp)
        self.opt = L.TransportOptions(int(upwind), int(nonOrth), nNonOrthCorr, maxIter, relTol,
                                      absTol)

    def gamma(self, k):
        g = np.empty(self.mesh.nFaces)
        check(lib().an_face_gamma_from_cells(self.mesh._h, ptr(f64(k)), ptr(g), L.AN_HOST))
        return g

    def solve(self, k, bc_vals):
        phi = np.empty(self.mesh.nCells)
        info = L.SolveInfo()
        st = lib().an_solve_steady_scalar_transport(
            self.mesh._h, ptr(self.flux), ptr(self.gamma(k)), ptr(self.Su), ptr(self.Sp),
            ptr(self.bc_types), ptr(f64(bc_vals)), C.byref(self.opt), None, ptr(phi), L.AN_HOST,
            C.byref(info))
        check(st, allow=(L.AN_ENOCONV,))
        return phi

    def adjoint(self, k, bc_vals, phi, dJdphi):
        """-> (dJ/dk [nCells], dJ/dbcVal [nBF], adjoint field [nCells])."""
        m = self.mesh
        dg = np.empty(m.nFaces)
        db = np.empty(m.nBoundaryFaces)
        psi = np.empty(m.nCells)
        info = L.SolveInfo()
        st = lib().an_transport_adjoint(m._h, ptr(self.flux), ptr(self.gamma(k)), ptr(self.Su),
                                        ptr(self.Sp), ptr(self.bc_types), ptr(f64(bc_vals)),
                                        C.byref(self.opt), ptr(f64(phi)), ptr(f64(dJdphi)),
                                        ptr(dg), ptr(db), ptr(psi), L.AN_HOST, C.byref(info))
        check(st, allow=(L.AN_ENOCONV,))
        dk = np.empty(m.nCells)
        check(lib().an_dgamma_to_cells(m._h, ptr(dg), ptr(dk), L.AN_HOST))
        return dk, db, psi


def lsq_gradient_adjoint(mesh, bc_types, bc_vals, w):
    """(d grad/d phi)^T w and (d grad/d bcVal)^T w for the LP gradient."""
    pb = np.empty(mesh.nCells)
    pbB = np.empty(mesh.nBoundaryFaces)
    check(lib().an_lsq_gradient_adjoint(mesh._h, ptr(i32(bc_types)), ptr(f64(bc_vals)),
                                        ptr(f64(np.asarray(w).reshape(-1))), ptr(pb), ptr(pbB),
                                        L.AN_HOST))
    return pb, pbB


def outlet_functional(mesh, patch_name, weight=None):
    """J(phi) = sum over the patch's faces of w_f A_f phi_P(f) (face value of a
    zero-gradient outlet = owner value).  Returns (J, dJ/dphi) callables."""
    k = mesh.patch_index(patch_name)
    s, n = mesh.patchStartFace[k], mesh.patchNFaces[k]
    owners = mesh.owner[s:s + n]
    wa = mesh.Af[s:s + n] * (np.ones(n) if weight is None else weight(mesh.xf[s:s + n]))
    grad = np.zeros(mesh.nCells)
    np.add.at(grad, owners, wa)
    return (lambda phi: float(np.dot(wa, phi[owners]))), (lambda phi: grad.copy())


def torch_transport_function():
    """torch.autograd.Function (k, bcVal) -> phi whose backward is the GPU
    adjoint; imported lazily so the library does not depend on torch."""
    import torch

    class TransportSolve(torch.autograd.Function):
        @staticmethod
        def forward(ctx, k, bc_vals, problem):
            kn = k.detach().cpu().double().numpy()
            bn = bc_vals.detach().cpu().double().numpy()
            phi = problem.solve(kn, bn)
            ctx.problem = problem
            ctx.save_for_backward(k, bc_vals)
            ctx.phi = phi
            return torch.from_numpy(phi).to(k.dtype)

        @staticmethod
        def backward(ctx, grad_phi):
            k, bc_vals = ctx.saved_tensors
            dk, db, _ = ctx.problem.adjoint(k.detach().cpu().double().numpy(),
                                            bc_vals.detach().cpu().double().numpy(), ctx.phi,
                                            grad_phi.detach().cpu().double().numpy())
            return torch.from_numpy(dk).to(k.dtype), torch.from_numpy(db).to(k.dtype), None

    return TransportSolve
