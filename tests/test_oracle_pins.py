"""Pins for the oracle functions that the cross-checks of test_oracle_ops.py
leave open (CPU only):

* momentum assembly at rho = 0 is the LP elliptic operator with gamma = mu
  (matrix, Dirichlet diagonal/rhs, non-orthogonal rhs) -- GSV1 main.cu:
  1804-1896 against LP/src/scalar_elliptic.cu:60-136;
* central convection on a uniform orthogonal box has the closed form
  +-F/2 per face and a zero diagonal in interior rows
  (LS/src/scalar_transport_library.cu:250-357; momentum at mu = 0 equals it);
* the Rhie-Chow predicted flux of a cubic pressure with exact cell
  gradients has the closed form rho A rAU a h^2 / 2, sign included
  (GSV1 main.cu:1474-1521);
* the SIMPLE driver (main.cu:3388-3634) converges on a plane-Poiseuille
  channel to the closed-form solution of the discrete fully developed
  problem (a tridiagonal system derived here from the finite-volume
  stencil), independently of the relaxation factors, and that discrete
  solution tends to the analytic parabola and dp/dx = -8 mu U / H^2 at
  second order;
* one SIMPLE outer iteration of the oracle equals an independent
  restatement of the driver order in Python that composes the pinned
  building blocks with scipy direct solves.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from anabasis_amd.libpoisson import Mesh
from oracle import pyoracle as O

from meshes import bc_arrays

PATCHES = ("xmin", "xmax", "ymin", "ymax", "zmin", "zmax")


def box(nx, ny, nz, lx=1.0, ly=1.0, lz=1.0, jitter=0.0, seed=0):
    m = Mesh.hexbox(nx, ny, nz, lx, ly, lz, jitter=jitter, seed=seed)
    assert tuple(m.patchNames) == PATCHES
    return m, O.OMesh.from_product(m)


def patch_types(dirichlet=()):
    return np.array([1 if p in dirichlet else 0 for p in PATCHES], np.int32)


def csr(om, vals):
    return sp.csr_matrix((vals, om.cols, om.rowOffsets), shape=(om.nCells, om.nCells))


# ------------------------------------------------------------------ (i)
def test_momentum_at_zero_density_is_the_elliptic_operator():
    """rho = 0 removes convection; what is left must be LP's diffusion
    operator with gamma = mu: alpha = mu A / (n.d) = mu |S|^2 / (d.S),
    t = mu (S - (A / n.d) d) = gamma T, the Dirichlet diagonal alpha and rhs
    alpha g + gamma T.grad(q_P), a source -dp/dx_i V = f V for
    gradPcomp = -f, and nothing on zero-gradient faces."""
    m, om = box(6, 5, 4, 1.0, 0.8, 0.7, jitter=0.25, seed=4)
    nC, nF, nIF = om.nCells, om.nFaces, om.nInternalFaces
    rng = np.random.default_rng(12)
    mu = 0.037
    dir_patches = ("xmin", "ymax", "zmin")
    tQ = patch_types(dir_patches)
    qB = rng.standard_normal(nF)
    gq = rng.standard_normal((nC, 3))
    f = rng.standard_normal(nC)
    uc = [rng.standard_normal(nC) for _ in range(3)]  # irrelevant at rho = 0
    vm, rm = O.momentum_assemble(om, gq, *uc, -f, tQ, qB, tQ, qB, tQ, qB, tQ, qB, 0.0, mu)
    t, v = bc_arrays(m, dirichlet=dir_patches)
    v[t == 1] = qB[nIF:][t == 1]
    ve, re = O.elliptic_assemble(om, np.full(nF, mu), f, t, v, gq, nonorth=True)
    assert np.abs(vm - ve).max() <= 1e-12 * np.abs(ve).max()
    assert np.abs(rm - re).max() <= 1e-12 * np.abs(re).max()
    # and the operator is not trivially zero in either part
    assert np.abs(ve).max() > 0.1 * mu and np.abs(re - f * om.vol).max() > 1e-3


# ----------------------------------------------------------------- (ii)
def test_central_convection_closed_form():
    """Uniform U on a uniform orthogonal box with cell sizes h = (hx, hy, hz):
    every internal face has lambda = 1/2 and F = U_k A_k, so an interior row
    reads A[c, c+e_k] = +U_k A_k / 2, A[c, c-e_k] = -U_k A_k / 2 and a zero
    diagonal.  A boundary cell on a zero-gradient (outflow) face adds F_b to
    its diagonal; on a Dirichlet face it moves -F_b q_B to the rhs."""
    nx, ny, nz, L = 6, 5, 4, (1.2, 0.8, 0.7)
    m, om = box(nx, ny, nz, *L)
    h = np.array(L) / np.array([nx, ny, nz])
    Ak = np.array([h[1] * h[2], h[0] * h[2], h[0] * h[1]])
    U = np.array([0.6, -0.3, 0.45])
    nC, nF, nIF = om.nCells, om.nFaces, om.nInternalFaces
    flux = om.Sf @ U
    tb = np.zeros(nF - nIF, np.int32)
    tb[om.bPatch[nIF:] == 1] = 1  # xmin Dirichlet, others zero-gradient
    qb = np.zeros(nF - nIF)
    qb[tb == 1] = 2.0
    vals, rhs = O.transport_assemble(om, flux, np.zeros(nF), np.zeros(nC), np.zeros(nC), tb, qb,
                                     np.zeros((nC, 3)), nonorth=False, upwind=False)
    A = csr(om, vals).toarray()
    stride = (1, nx, nx * ny)

    def cell(i, j, k):
        return i + nx * (j + ny * k)

    for (i, j, k) in ((2, 2, 1), (3, 1, 2), (1, 3, 2)):
        c = cell(i, j, k)
        expect = np.zeros(nC)
        for d in range(3):
            expect[c + stride[d]] += U[d] * Ak[d] / 2
            expect[c - stride[d]] -= U[d] * Ak[d] / 2
        assert np.allclose(A[c], expect, rtol=1e-13, atol=1e-16), (i, j, k)
        assert abs(A[c, c]) < 1e-16
    # boundary rows: the xmax cell (2, 2, 1) -> (5, 2, 1) adds the outflow
    # F_b = U_x A_x on its diagonal; the xmin cell moves -F_b q_B = +U_x A_x 2
    c = cell(nx - 1, 2, 1)
    assert np.isclose(A[c, c], U[0] * Ak[0] / 2, rtol=1e-13)  # F_b - F_face/2
    c = cell(0, 2, 1)
    assert np.isclose(rhs[c], U[0] * Ak[0] * 2.0, rtol=1e-13)
    # GSV1 momentum with rho = 1, mu = 0 and the same uniform velocity
    # assembles the same matrix (central convection of the interpolated
    # old velocity, main.cu:1839-1846)
    tU = patch_types(("xmin",))
    qB = np.zeros(nF)
    qB[om.bPatch == 1] = 2.0
    uB = [np.zeros(nF) for _ in range(3)]
    for d in range(3):
        uB[d][om.bPatch == 1] = U[d]
    vm, rm = O.momentum_assemble(om, np.zeros((nC, 3)), *(np.full(nC, U[d]) for d in range(3)),
                                 np.zeros(nC), tU, qB, tU, uB[0], tU, uB[1], tU, uB[2], 1.0, 0.0)
    assert np.abs(vm - vals).max() < 1e-14
    assert np.abs(rm - rhs).max() < 1e-14


# ---------------------------------------------------------------- (iii)
def test_rhiechow_closed_forms():
    """Uniform orthogonal box, uniform U and rAU = r.
    (a) zero cell gradients: phi* = rho A U.n - rho A r (p_N - p_P) / h_k;
    (b) p = sum_k a_k x_k^3 with exact cell gradients 3 a_k x_k^2: the
        face-interpolated gradient over-estimates the jump by a_k h_k^3 / 2,
        so phi* = rho A U.n + rho A r a_k h_k^2 / 2 (positive for a_k > 0)."""
    nx, ny, nz, L = 6, 5, 4, (1.2, 0.8, 0.7)
    m, om = box(nx, ny, nz, *L)
    h = np.array(L) / np.array([nx, ny, nz])
    nC, nF, nIF = om.nCells, om.nFaces, om.nInternalFaces
    U = np.array([0.4, 0.1, -0.25])
    rho, r = 1.3, 0.37
    a = np.array([0.8, -1.1, 0.6])
    p = (om.cc ** 3) @ a
    t = patch_types(("xmin",))
    ub = [np.zeros(nF) for _ in range(3)]
    uc = [np.full(nC, U[d]) for d in range(3)]
    P, N = om.owner[:nIF], om.neigh
    k = np.argmax(np.abs(om.nf[:nIF]), axis=1)  # face direction
    # each internal face is normal to one axis, oriented from P to N
    assert np.all(np.abs(om.nf[np.arange(nIF), k] - 1.0) <= 1e-14)
    A = om.Af[:nIF]
    base = rho * A * (om.nf[:nIF] @ U)
    # (a)
    phi = O.rhiechow_flux(om, *uc, p, np.zeros((nC, 3)), np.full(nC, r), t, ub[0], t, ub[1], t,
                          ub[2], rho)
    expect = base - rho * A * r * (p[N] - p[P]) / h[k]
    assert np.abs(phi[:nIF] - expect).max() <= 1e-12 * np.abs(expect).max()
    # (b)
    g = 3.0 * om.cc ** 2 * a
    phi = O.rhiechow_flux(om, *uc, p, g, np.full(nC, r), t, ub[0], t, ub[1], t, ub[2], rho)
    expect = base + rho * A * r * a[k] * h[k] ** 2 / 2.0
    assert np.abs(phi[:nIF] - expect).max() <= 1e-12 * np.abs(base).max()
    assert np.all(np.sign(phi[:nIF] - base) == np.sign(a[k]))


# ----------------------------------------------------------------- (iv)
def discrete_poiseuille(ny, H, U, mu):
    """Fully developed flow between no-slip walls y = 0, H on ny uniform
    cells: the finite-volume stencil reads mu (u_{j+1} - 2 u_j + u_{j-1}) / h^2
    = G in interior cells and mu (u_1 - 3 u_0) / h^2 = G at a wall (wall
    distance h/2); the flow rate is the one the inlet faces carry, the
    midpoint sum of 4 U y (H - y) / H^2.  Returns (u_j, G = dp/dx)."""
    h = H / ny
    y = (np.arange(ny) + 0.5) * h
    A = (np.diag(np.full(ny, -2.0)) + np.diag(np.ones(ny - 1), 1) +
         np.diag(np.ones(ny - 1), -1))
    A[0, 0] = A[-1, -1] = -3.0
    shape = np.linalg.solve(A / h ** 2, np.ones(ny))  # u for G / mu = 1
    Q = h * np.sum(4 * U * y * (H - y) / H ** 2)
    scale = Q / (h * shape.sum())
    return shape * scale, mu * scale


def poiseuille_run(ny, uRelax=0.5, pRelax=0.3, L=6.0, H=1.0, W=0.25, U=1.0, mu=1.0):
    """Channel [0,L]x[0,H]x[0,W] at Re = rho U H / mu = 1: no-slip y walls,
    zero-gradient z walls (one cell across z), p = 0 at x = L with a
    zero-gradient velocity, and an inlet at x = 0 that carries the discrete
    fully developed profile (inlet faces sit at the cells' y).  The inlet
    cells still see a disturbance: the GSV1 least-squares gradient counts the
    zero-gradient inlet face as a zero difference (main.cu:1403-1440), so
    their pressure gradient is about half of G.  It decays downstream; for
    x in [L/2, L - H/2] the field must be the x-invariant discrete solution:
    u = u_j, v = w = 0 and p = G (x - L)."""
    nx = int(round(L / H * ny))
    m, om = box(nx, ny, 1, L, H, W)
    nC, nF = om.nCells, om.nFaces
    ud, Gd = discrete_poiseuille(ny, H, U, mu)
    tU = patch_types(("xmin", "ymin", "ymax"))
    tP = patch_types(("xmax",))
    ub = [np.zeros(nF) for _ in range(3)]
    sel = om.bPatch == 1
    j = np.floor(om.xf[sel, 1] / (H / ny)).astype(int)
    ub[0][sel] = ud[j]
    amg = O.amg_options(maxIter=1000, relTol=0.0, absTol=1e-13)
    par = O.SimpleParams(1.0, mu, 1, 0, uRelax, pRelax, 1e-10, 1e-11, 1e-11, 1000, 1e-13, 0.0,
                         1000, 1e-13, 0.0, 10, amg)
    r = O.simple_run(om, par, tU, tP, *ub, np.zeros(nF), 3000, np.full(nC, ub[0][sel].mean()),
                     np.zeros(nC), np.zeros(nC), np.zeros(nC))
    assert 0 < r["done"] < 3000, "did not reach the stop test"
    xc = om.cc[:nx, 0]
    cols = (xc >= 0.5 * L) & (xc <= L - 0.5 * H)
    u = r["u"].reshape(ny, nx)[:, cols]
    v = r["v"].reshape(ny, nx)[:, cols]
    p = r["p"].reshape(ny, nx)[:, cols]
    p_exact = Gd * (xc[cols] - L)  # p = 0 on the outlet face
    return u, v, p, p_exact, ud, Gd, r


def test_simple_plane_poiseuille_matches_discrete_closed_form():
    U, mu, H = 1.0, 1.0, 1.0
    err_u, err_g = [], []
    for ny in (4, 8, 16):
        u, v, p, pe, ud, Gd, _ = poiseuille_run(ny, U=U, mu=mu, H=H)
        # the converged state is the discrete fully developed solution
        assert np.abs(u - ud[:, None]).max() <= 1e-8 * U, ny
        assert np.abs(v).max() <= 1e-8 * U, ny
        assert np.abs(p - pe[None, :]).max() <= 1e-8 * np.abs(pe).max(), ny
        y = (np.arange(ny) + 0.5) * H / ny
        err_u.append(np.abs(ud - 4 * U * y * (H - y) / H ** 2).max())
        err_g.append(abs(Gd + 8 * mu * U / H ** 2))
    # the discrete solution converges to the analytic one: dp/dx at second
    # order; the velocity's max error (set by the wall cells) approaches it
    og = np.log2(np.array(err_g[:-1]) / np.array(err_g[1:]))
    ou = np.log2(np.array(err_u[:-1]) / np.array(err_u[1:]))
    assert np.all(og > 1.8), og
    assert ou[-1] > 1.5 and ou[-1] > ou[0], ou


def test_simple_poiseuille_state_does_not_depend_on_relaxation():
    u1, _, p1, _, _, _, r1 = poiseuille_run(8, 0.5, 0.3)
    u2, _, p2, _, _, _, r2 = poiseuille_run(8, 0.3, 0.15)
    assert r1["done"] != r2["done"]  # a different path ...
    assert np.abs(u1 - u2).max() <= 1e-9  # ... to the same fixed point
    assert np.abs(p1 - p2).max() <= 1e-9 * np.abs(p1).max()


# ------------------------------------------------------------------ (v)
def _relchg(a, b):
    return np.sqrt(np.sum((a - b) ** 2) / (1.0 + np.sum(a * a)))


def simple_step_restated(om, rho, mu, nVel, nNonOrth, uRelax, pRelax, corrTol, tU, tP, fb, pb,
                         u, v, w, p):
    """One outer iteration in the order of GSV1 main.cu:3388-3603 and the
    component loop :3350-3384, built from the oracle's operators with direct
    sparse solves in place of BiCGSTAB and PCG."""
    nIF = om.nInternalFaces
    bpatch = om.bPatch[nIF:] - 1
    tUb, tPb = tU[bpatch], tP[bpatch]
    old = [u.copy(), v.copy(), w.copy()]
    pOld = p.copy()
    grad = lambda q, t, vals: O.lsq_gradient(om, q, t, vals[nIF:], mode="gsv1")
    gradP = grad(pOld, tPb, pb)
    star = []
    A = vals = rAU = None
    for comp in range(3):
        q = old[comp].copy()
        for it in range(1, nVel + 1):
            gq = grad(q, tUb, fb[comp])
            if comp == 0 and it == 1:
                vals, rhs = O.momentum_assemble(om, gq, *old, gradP[:, comp], tU, fb[comp], tU,
                                                fb[0], tU, fb[1], tU, fb[2], rho, mu)
                vals, rhs = O.momentum_relax(om, vals, rhs, q, uRelax)
                rAU = om.vol / vals[om.diagPos]
                A = csr(om, vals).tocsc()
            else:
                rhs = O.momentum_rhs_only(om, gq, *old, gradP[:, comp], tU, fb[comp], tU, fb[0],
                                          tU, fb[1], tU, fb[2], rho, mu)
                rhs = O.momentum_relax_rhs_only(om, vals, rhs, q, uRelax)
            qn = spla.spsolve(A, rhs)
            rc = _relchg(qn, q)
            q = qn
            if rc < corrTol:
                break
        star.append(q)
    Ap = csr(om, O.pressure_assemble(om, rAU, tP)).tocsc()
    phiStar = O.rhiechow_flux(om, *star, pOld, gradP, rAU, tU, fb[0], tU, fb[1], tU, fb[2], rho)
    divStar = O.continuity(om, phiStar)
    pc = np.zeros(om.nCells)
    for k in range(nNonOrth + 1):
        _, divNo = O.pressure_nonorth(om, rAU, tP, grad(pc, tPb, pb))
        pc = spla.spsolve(Ap, -divStar + divNo)
    phiNo, _ = O.pressure_nonorth(om, rAU, tP, grad(pc, tPb, pb))
    phi = O.correct_flux(om, rAU, tP, pb, phiStar, pc, phiNo)
    mass = np.abs(O.continuity(om, phi)).max()
    p = pOld + pRelax * pc
    gc = grad(pc, tPb, pb)
    u, v, w = (star[d] - rAU * gc[:, d] for d in range(3))
    return u, v, w, p, phi, mass


def test_simple_outer_iteration_matches_independent_restatement():
    m, om = box(8, 4, 3, 2.0, 1.0, 0.8, jitter=0.2, seed=21)
    nC, nF = om.nCells, om.nFaces
    rho, mu, nVel, nNo, ur, pr = 1.0, 0.05, 2, 2, 0.6, 0.35
    tU = patch_types(("xmin", "ymin", "ymax", "zmin", "zmax"))
    tP = patch_types(("xmax",))
    fb = [np.zeros(nF) for _ in range(3)]
    sel = om.bPatch == 1
    y, z = om.xf[sel, 1], om.xf[sel, 2]
    fb[0][sel] = 16 * 0.45 * y * z * (1 - y) * (0.8 - z) / 0.64
    fb[1][sel] = 0.02 * y
    pb = np.zeros(nF)
    state = [np.full(nC, fb[0][sel].mean()), np.full(nC, fb[1][sel].mean()), np.zeros(nC),
             np.zeros(nC)]
    amg = O.amg_options(maxIter=2000, relTol=0.0, absTol=1e-15)
    # corrTol 0: every velocity pass runs in both restatements
    par = O.SimpleParams(rho, mu, nVel - 1, nNo, ur, pr, 0.0, -1.0, -1.0, 4000, 1e-15, 0.0, 2000,
                         1e-15, 0.0, 10, amg)
    ref = O.simple_run(om, par, tU, tP, *fb, pb, 3, *state)
    assert ref["done"] == 3
    u, v, w, p = state
    for it in range(3):
        u, v, w, p, phi, mass = simple_step_restated(om, rho, mu, nVel, nNo, ur, pr, 0.0, tU, tP,
                                                     fb, pb, u, v, w, p)
        assert np.isclose(mass, ref["stats"][it].massRes, rtol=1e-6, atol=1e-14)
    for k, a in (("u", u), ("v", v), ("w", w), ("p", p), ("phi", phi)):
        b = ref[k]
        assert np.abs(a - b).max() <= 1e-9 * np.abs(b).max(), k
    # the restatement is sensitive to the driver order: dropping the
    # pressure relaxation or correcting u with the gradient of p instead of
    # p' gives a visibly different state after one iteration
    u1, v1, w1, p1, _, _ = simple_step_restated(om, rho, mu, nVel, nNo, ur, 1.0, 0.0, tU, tP, fb,
                                                pb, *state)
    r1 = O.simple_run(om, par, tU, tP, *fb, pb, 1, *state)
    assert np.abs(p1 - r1["p"]).max() > 1e-3 * np.abs(r1["p"]).max()


# --------------------------------------------------------- PCG breakdown
def test_pcg_breakdown_stops_before_the_update():
    """A symmetric indefinite matrix (Laplacian shifted between two of its
    eigenvalues): the PCG reports a breakdown (p.q <= 0) and returns the last
    good iterate, finite; when the very first p.q is negative x stays x0
    (HYPRE's PCG checks gamma = p.Ap before updating x)."""
    m, om = box(10, 9, 8, jitter=0.2, seed=3)
    t, v = bc_arrays(m, dirichlet=("xmin",))
    vals, _ = O.elliptic_assemble(om, np.ones(om.nFaces), np.zeros(om.nCells), t, v,
                                  np.zeros((om.nCells, 3)))
    ev = np.linalg.eigvalsh(csr(om, vals).toarray())
    b = np.random.default_rng(1).standard_normal(om.nCells)
    for k, first in ((3, False), (100, True)):
        shifted = vals.copy()
        shifted[om.diagPos] -= 0.5 * (ev[k] + ev[k + 1])
        x0 = np.zeros(om.nCells)
        x, it, _, _ = O.pcg_amg(om.rowOffsets, om.cols, shifted, b, x0=x0, relTol=1e-10,
                                absTol=0.0, coarseSize=32)
        assert it < 0 and np.all(np.isfinite(x))
        assert np.all(x == 0.0) == first
