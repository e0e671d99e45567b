"""GPU parity of the HIP kernels against the CPU oracle (through the C ABI).

Integer addressing and AMG aggregation must be bit-exact; the assembly,
gradient, SpMV and V-cycle kernels evaluate the reference's expressions in
the reference's order (no FMA contraction), so they are required to match
the serial oracle bit for bit as well.  Solver results are compared at
converged-solution level (HYPRE-style histories are not pinnable).
"""
import zlib

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from anabasis_amd import _lib as L
from anabasis_amd import libpoisson as lp
from oracle import pyoracle as O

from meshes import bc_arrays

pytestmark = pytest.mark.gpu


def crc(arrs):
    c = 0
    for a in arrs:
        c = zlib.crc32(np.ascontiguousarray(a, np.int32).tobytes(), c)
    return c


@pytest.fixture(scope="module")
def meshes(gpu):
    return {
        "hex": lp.Mesh.hexbox(11, 9, 8, 1.0, 0.9, 0.8, jitter=0.2, seed=11),
        "tet": lp.Mesh.tetcube(5),
        "cyl": lp.Mesh.cylinder(9000),
    }


@pytest.fixture(scope="module")
def omeshes(meshes):
    return {k: O.OMesh.from_product(m) for k, m in meshes.items()}


# ---------------------------------------------------------- addressing
def test_csr_cylinder_fixture_bitexact(gpu, cylinder_fixture):
    fx = cylinder_fixture
    pat = lp.csr_from_faces(fx["nCells"], fx["owner"], fx["neigh"])
    assert pat.nnz == fx["nnz"] == 617000
    arrs = (pat.rowOffsets, pat.cols, pat.diagPos, pat.facePP, pat.facePN, pat.faceNP, pat.faceNN)
    assert crc(arrs) == fx["csr_crc"]
    ref = O.pattern(fx["nCells"], fx["owner"], fx["neigh"])
    for a, b in zip(arrs, ref):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("kind", ["hex", "tet", "cyl"])
def test_csr_generated_bitexact(meshes, omeshes, kind):
    m, om = meshes[kind], omeshes[kind]
    pat = lp.build_scalar_pattern(m)
    for a, b in zip((pat.rowOffsets, pat.cols, pat.diagPos, pat.facePP, pat.facePN, pat.faceNP,
                     pat.faceNN), (om.rowOffsets, om.cols, om.diagPos, om.PP, om.PN, om.NP, om.NN)):
        assert np.array_equal(a, b)


def test_csr_duplicate_faces_and_empty_boundary(gpu):
    # two faces between cells 0 and 1 (polyhedral meshes can have this), a
    # cell with no boundary faces and an isolated boundary-only cell
    owner = np.array([0, 0, 1, 0, 2, 3], np.int32)
    neigh = np.array([1, 1, 2], np.int32)
    pat = lp.csr_from_faces(4, owner, neigh)
    ref = O.pattern(4, owner, neigh)
    for a, b in zip((pat.rowOffsets, pat.cols, pat.diagPos, pat.facePP, pat.facePN, pat.faceNP,
                     pat.faceNN), ref):
        assert np.array_equal(a, b)
    assert pat.nnz == 4 + 2 * 2


def test_csr_rejects_bad_indices(gpu):
    with pytest.raises(L.AnError):
        lp.csr_from_faces(3, np.array([0, 5, 1], np.int32), np.array([1], np.int32))


# ------------------------------------------------------------ gradients
@pytest.mark.parametrize("kind", ["hex", "tet", "cyl"])
@pytest.mark.parametrize("mode", [L.AN_GRAD_GSV1, L.AN_GRAD_LP])
def test_lsq_gradient_bitexact(meshes, omeshes, kind, mode):
    m, om = meshes[kind], omeshes[kind]
    rng = np.random.default_rng(5)
    phi = rng.standard_normal(m.nCells)
    nb = m.nBoundaryFaces
    t = (rng.random(nb) < 0.5).astype(np.int32)
    v = rng.standard_normal(nb)
    g = lp.compute_lsq_gradient(m, phi, t, v, mode)
    ref = O.lsq_gradient(om, phi, t, v, "gsv1" if mode == L.AN_GRAD_GSV1 else "lp")
    assert np.array_equal(g, ref)


# ------------------------------------------------------------- assembly
@pytest.mark.parametrize("kind", ["hex", "tet", "cyl"])
def test_elliptic_assembly_bitexact(meshes, omeshes, kind):
    m, om = meshes[kind], omeshes[kind]
    rng = np.random.default_rng(9)
    gam = 0.5 + rng.random(m.nFaces)
    src = rng.standard_normal(m.nCells)
    nb = m.nBoundaryFaces
    t = (rng.random(nb) < 0.6).astype(np.int32)
    v = rng.standard_normal(nb)
    grad = rng.standard_normal((m.nCells, 3))
    for nonorth, use_ref in ((True, False), (False, True)):
        vals, rhs = lp.assemble_scalar_elliptic_system(m, gam, src, t, v, grad, nonorth, use_ref,
                                                       3, 1.25, nnz=om.nnz)
        rv, rr = O.elliptic_assemble(om, gam, src, t, v, grad, nonorth, use_ref, 3, 1.25)
        assert np.array_equal(vals, rv) and np.array_equal(rhs, rr)


@pytest.mark.parametrize("upwind", [False, True])
def test_transport_assembly_bitexact(meshes, omeshes, upwind):
    m, om = meshes["cyl"], omeshes["cyl"]
    rng = np.random.default_rng(2)
    flux = rng.standard_normal(m.nFaces)
    gam = 0.1 + rng.random(m.nFaces)
    Su, Sp = rng.standard_normal(m.nCells), -rng.random(m.nCells)
    nb = m.nBoundaryFaces
    t = rng.integers(0, 3, nb).astype(np.int32)
    v = rng.standard_normal(nb)
    grad = rng.standard_normal((m.nCells, 3))
    vals, rhs = lp.assemble_scalar_transport_system(m, flux, gam, Su, Sp, t, v, grad, True, upwind,
                                                    nnz=om.nnz)
    rv, rr = O.transport_assemble(om, flux, gam, Su, Sp, t, v, grad, True, upwind)
    assert np.array_equal(vals, rv) and np.array_equal(rhs, rr)


# -------------------------------------------------------- SpMV / AMG / PCG
def laplacian(om, m, dirichlet=("xmin",)):
    t, v = bc_arrays(m, dirichlet=dirichlet)
    vals, _ = O.elliptic_assemble(om, np.ones(om.nFaces), np.zeros(om.nCells), t, v,
                                  np.zeros((om.nCells, 3)))
    return vals


@pytest.mark.parametrize("kind", ["hex", "cyl"])
def test_spmv_amg_vcycle_bitexact(meshes, omeshes, kind):
    m, om = meshes[kind], omeshes[kind]
    dirich = ("xmin",) if kind == "hex" else ("patch_2_0",)
    vals = laplacian(om, m, dirich)
    S = lp.PCGSolver(om.rowOffsets, om.cols, coarseSize=64, kcycle=0)
    S.set_values(vals)
    x = np.random.default_rng(1).standard_normal(om.nCells)
    assert np.array_equal(S.spmv(x), O.spmv(om.rowOffsets, om.cols, vals, x))
    agg, nc = S.aggregates()
    ragg, rnc = O.amg_aggregate(om.rowOffsets, om.cols, vals, coarseSize=64)
    assert nc == rnc and np.array_equal(agg, ragg)
    # whole hierarchy (Galerkin sums, smoother weights, dense coarse inverse)
    z = S.amg_apply(x)
    zo, nl = O.amg_vcycle(om.rowOffsets, om.cols, vals, x, coarseSize=64, kcycle=0)
    assert nl >= 3 and np.array_equal(z, zo)
    for sweeps in (2, 3):  # multi-sweep smoothing (buffer ping-pong on every level)
        S2 = lp.PCGSolver(om.rowOffsets, om.cols, coarseSize=64, sweeps=sweeps, kcycle=0)
        S2.set_values(vals)
        z2 = S2.amg_apply(x)
        zo2, _ = O.amg_vcycle(om.rowOffsets, om.cols, vals, x, coarseSize=64, sweeps=sweeps,
                              kcycle=0)
        assert np.array_equal(z2, zo2), sweeps


@pytest.mark.parametrize("kind", ["hex", "cyl"])
def test_chebyshev_vcycle_bitexact(meshes, omeshes, kind):
    """The degree-2 Chebyshev smoother (Gershgorin bound, coefficients and
    the scaled diagonal computed on the GPU): the oracle's V-cycle bit for
    bit on the finest level only, on every smoothed level, in FP64 and --
    to FP32 rounding -- in the FP32 cycle of precision 2."""
    m, om = meshes[kind], omeshes[kind]
    dirich = ("xmin",) if kind == "hex" else ("patch_2_0",)
    vals = laplacian(om, m, dirich)
    x = np.random.default_rng(3).standard_normal(om.nCells)
    for levels in (1, 0):
        kw = dict(coarseSize=64, kcycle=0, smoother=1, chebLevels=levels)
        S = lp.PCGSolver(om.rowOffsets, om.cols, relTol=1e-10, absTol=0.0, **kw)
        S.set_values(vals)
        z = S.amg_apply(x)
        zo, nl = O.amg_vcycle(om.rowOffsets, om.cols, vals, x, **kw)
        assert nl >= 3 and np.array_equal(z, zo), levels
        zj, _ = O.amg_vcycle(om.rowOffsets, om.cols, vals, x, coarseSize=64, kcycle=0)
        assert np.linalg.norm(zj - z) > 1e-4 * np.linalg.norm(z)
        # new values without a rebuild: the finest level's bound and
        # coefficients follow them (the oracle's stale-hierarchy refresh)
        S.set_values(2.0 * vals, rebuild=False)
        h = O.pcg_create(om.rowOffsets, om.cols, **kw)
        try:
            O.pcg_set_values(h, vals, rebuild=True)
            O.pcg_set_values(h, 2.0 * vals, rebuild=False)
            xs = S.solve(x)[0]
            xo, ito, _ = O.pcg_solve(h, x, relTol=1e-10, absTol=0.0, maxIter=500)
            assert ito > 0
        finally:
            O.pcg_destroy(h)
        assert np.linalg.norm(xs - xo) <= 1e-8 * np.linalg.norm(xo)


def test_chebyshev_pcg_matches_oracle(meshes, omeshes):
    """PCG + K-cycle AMG with Chebyshev smoothing on every smoothed level:
    the oracle's iteration count and residual history (1e-8), FP64 and
    precision 2."""
    m, om = meshes["cyl"], omeshes["cyl"]
    vals = laplacian(om, m, ("patch_2_0",))
    b = np.random.default_rng(43).standard_normal(om.nCells)
    kw = dict(coarseSize=64, kcycle=1, smoother=1, chebLevels=0)
    xo, it, rr, ho = O.pcg_amg(om.rowOffsets, om.cols, vals, b, relTol=1e-10, absTol=0.0,
                               hist_cap=500, **kw)
    ho = ho[: it + 1]
    for precision in (0, 2):
        S = lp.PCGSolver(om.rowOffsets, om.cols, relTol=1e-10, absTol=0.0, precision=precision,
                         **kw)
        S.set_values(vals)
        x, info = S.solve(b)
        assert info.converged and np.linalg.norm(x - xo) <= 1e-8 * np.linalg.norm(xo)
        if precision == 0:
            h = S.residual_history()
            assert info.iterations == it and np.all(np.abs(h - ho) <= 1e-8 * ho)
        else:
            assert abs(info.iterations - it) <= 2


@pytest.mark.parametrize("kind", ["hex", "cyl"])
def test_kcycle_matches_oracle(meshes, omeshes, kind):
    """K-cycle preconditioner: the oracle's cycle; its inner products are
    two-stage block sums (another summation order than the oracle's serial
    dot), so the two agree to rounding, and the GPU result is reproducible
    bit for bit."""
    m, om = meshes[kind], omeshes[kind]
    dirich = ("xmin",) if kind == "hex" else ("patch_2_0",)
    vals = laplacian(om, m, dirich)
    x = np.random.default_rng(2).standard_normal(om.nCells)
    for kc in ((1,) if kind == "hex" else (1, 2)):
        for sweeps in (1, 2):
            S = lp.PCGSolver(om.rowOffsets, om.cols, coarseSize=64, kcycle=kc, sweeps=sweeps)
            S.set_values(vals)
            z = S.amg_apply(x)
            assert np.array_equal(z, S.amg_apply(x))
            zo, nl = O.amg_vcycle(om.rowOffsets, om.cols, vals, x, coarseSize=64, kcycle=kc,
                                  sweeps=sweeps)
            zv, _ = O.amg_vcycle(om.rowOffsets, om.cols, vals, x, coarseSize=64, sweeps=sweeps,
                                 kcycle=0)
            assert nl >= kc + 2
            assert np.linalg.norm(z - zo) <= 1e-12 * np.linalg.norm(zo), (kc, sweeps)
            assert np.linalg.norm(zv - zo) > 1e-4 * np.linalg.norm(zo)


def test_pcg_kcycle_flexible_converges(meshes, omeshes):
    m, om = meshes["cyl"], omeshes["cyl"]
    vals = laplacian(om, m, ("patch_2_0",))
    b = np.random.default_rng(8).standard_normal(om.nCells)
    A = sp.csr_matrix((vals, om.cols, om.rowOffsets))
    xd = spla.spsolve(A.tocsc(), b)
    xo, ito, rr, _ = O.pcg_amg(om.rowOffsets, om.cols, vals, b, relTol=1e-12, absTol=0.0,
                               coarseSize=64, kcycle=2)
    V = lp.PCGSolver(om.rowOffsets, om.cols, relTol=1e-12, absTol=0.0, coarseSize=64, kcycle=0)
    V.set_values(vals)
    _, iv = V.solve(b)
    for precision in (0, 1):
        S = lp.PCGSolver(om.rowOffsets, om.cols, relTol=1e-12, absTol=0.0, coarseSize=64,
                         kcycle=2, precision=precision)
        S.set_values(vals)
        x, info = S.solve(b)
        assert info.converged and info.finalRelResNorm <= 1e-12
        assert np.linalg.norm(x - xd) <= 1e-9 * np.linalg.norm(xd)
        if precision == 0:
            assert abs(info.iterations - ito) <= 2
            assert info.iterations < iv.iterations
            x2, _ = S.solve(b)
            assert np.array_equal(x, x2)


def test_pcg_graph_replay_matches_eager(meshes, omeshes):
    """The preconditioner replayed from captured HIP graphs gives the eager
    launch sequence's results bit for bit, over several solves (host work in
    between) and after a hierarchy rebuild (graphs re-captured)."""
    m, om = meshes["cyl"], omeshes["cyl"]
    vals = laplacian(om, m, ("patch_2_0",))
    rng = np.random.default_rng(9)
    for precision in (0, 1):
        kw = dict(relTol=1e-10, absTol=0.0, coarseSize=64, precision=precision)
        E = lp.PCGSolver(om.rowOffsets, om.cols, graphs=0, **kw)
        G = lp.PCGSolver(om.rowOffsets, om.cols, graphs=1, **kw)
        for scale in (1.0, 2.0):
            E.set_values(scale * vals, rebuild=True)
            G.set_values(scale * vals, rebuild=True)
            for _ in range(3):
                b = rng.standard_normal(om.nCells)
                xe, ie = E.solve(b)
                xg, ig = G.solve(b)
                assert ie.converged and ie.iterations == ig.iterations
                assert np.array_equal(xe, xg), (precision, scale)


@pytest.mark.parametrize("kind", ["tet", "cyl"])
def test_pcg_residual_history_matches_oracle(meshes, omeshes, kind):
    """The per-iteration residual norms of a solve (what the reference's
    `monitor` option prints) follow the oracle's PCG + AMG iteration by
    iteration: same count, each norm within 1e-8 relative (only the dot
    products' summation order differs)."""
    m, om = meshes[kind], omeshes[kind]
    vals = laplacian(om, m, ("xmin",) if kind == "tet" else ("patch_2_0",))
    b = np.random.default_rng(41).standard_normal(om.nCells)
    S = lp.PCGSolver(om.rowOffsets, om.cols, relTol=1e-10, absTol=0.0, coarseSize=64)
    S.set_values(vals)
    x, info = S.solve(b)
    h = S.residual_history()
    xo, it, rr, ho = O.pcg_amg(om.rowOffsets, om.cols, vals, b, relTol=1e-10, absTol=0.0,
                               coarseSize=64, hist_cap=500)
    assert info.converged and info.iterations == it and len(h) == it + 1
    ho = ho[: it + 1]
    assert np.all(np.abs(h - ho) <= 1e-8 * ho), np.abs(h - ho) / ho


def test_pcg_repeatable(meshes, omeshes):
    """Two solvers fed the same sequence of solves (projection on, so the
    host reads scalars through both the per-iteration publication and the
    per-solve copies) return the same bits and iteration counts."""
    m, om = meshes["cyl"], omeshes["cyl"]
    vals = laplacian(om, m, ("patch_2_0",))
    for precision in (0, 1):
        kw = dict(relTol=1e-9, absTol=0.0, coarseSize=64, precision=precision, rhsProjection=4)
        S = [lp.PCGSolver(om.rowOffsets, om.cols, **kw) for _ in range(2)]
        for P in S:
            P.set_values(vals, rebuild=True)
        rng = np.random.default_rng(21)
        b0 = rng.standard_normal(om.nCells)
        for k in range(8):
            b = b0 + 0.1 * k * rng.standard_normal(om.nCells)
            (x1, i1), (x2, i2) = (P.solve(b.copy()) for P in S)
            assert i1.converged and i1.iterations == i2.iterations, (precision, k)
            assert np.array_equal(x1, x2), (precision, k)


def test_pcg_rhs_projection_many_vectors(gpu):
    """The most stored solutions (16) on a 110,592-row matrix: the projection's
    block partials (blocks x 16 components) stay inside the reducer, and every
    solve of a sequence still meets its tolerance on the true residual."""
    m = lp.Mesh.hexbox(48, 48, 48, jitter=0.2, seed=5)
    om = O.OMesh.from_product(m)
    vals = laplacian(om, m, ("xmin",))
    A = sp.csr_matrix((vals, om.cols, om.rowOffsets))
    P = lp.PCGSolver(om.rowOffsets, om.cols, relTol=0.0, absTol=1e-9, rhsProjection=16)
    P.set_values(vals)
    rng = np.random.default_rng(31)
    b0 = rng.standard_normal(om.nCells)
    for k in range(18):
        b = b0 + 0.05 * rng.standard_normal(om.nCells)
        x, info = P.solve(b)
        assert info.converged and np.linalg.norm(b - A @ x) <= 1.5e-9, k


def test_pcg_rhs_projection(meshes, omeshes):
    """Successive right-hand sides with one matrix: every solve still meets
    the tolerance (true residual) and agrees with the plain solve; a repeated
    right-hand side needs no iteration, nearby ones fewer; new values clear
    the stored solutions."""
    m, om = meshes["cyl"], omeshes["cyl"]
    vals = laplacian(om, m, ("patch_2_0",))
    A = sp.csr_matrix((vals, om.cols, om.rowOffsets))
    rng = np.random.default_rng(11)
    b1 = rng.standard_normal(om.nCells)
    kw = dict(relTol=0.0, absTol=1e-10, coarseSize=64)
    P = lp.PCGSolver(om.rowOffsets, om.cols, rhsProjection=8, **kw)
    Q = lp.PCGSolver(om.rowOffsets, om.cols, **kw)
    P.set_values(vals)
    Q.set_values(vals)
    x1, i1 = P.solve(b1)
    assert i1.converged and np.linalg.norm(b1 - A @ x1) <= 1.5e-10
    x2, i2 = P.solve(b1.copy())
    assert i2.iterations <= 1 and np.linalg.norm(b1 - A @ x2) <= 1.5e-10
    # right-hand sides of a contracting fixed-point iteration
    # x_k = A^-1 (b1 + N x_{k-1}), as in the non-orthogonal correction passes
    # (here N = A S, S = diag(0.5 +- 0.1), so A^-1 N = S)
    N = A @ sp.diags(0.5 + 0.1 * (rng.random(om.nCells) - 0.5))
    xp = xq = x1
    itp = itq = 0
    for k in range(6):
        b = b1 + N @ xq
        xp, ip = P.solve(b, x0=xp)
        xq, iq = Q.solve(b, x0=xq)
        assert ip.converged and np.linalg.norm(b - A @ xp) <= 1.5e-10
        assert np.linalg.norm(xp - xq) <= 1e-8 * np.linalg.norm(xq)
        itp += ip.iterations
        itq += iq.iterations
    assert itp < itq, (itp, itq)
    P.set_values(2.0 * vals)  # new matrix: the set is cleared
    x3, i3 = P.solve(b1)
    assert np.linalg.norm(b1 - 2.0 * (A @ x3)) <= 1.5e-10 and i3.iterations >= i1.iterations - 2


def test_pcg_amg_converges_to_direct_solution(meshes, omeshes):
    m, om = meshes["hex"], omeshes["hex"]
    vals = laplacian(om, m)
    b = np.random.default_rng(3).standard_normal(om.nCells)
    b -= b.mean()
    S = lp.PCGSolver(om.rowOffsets, om.cols, relTol=1e-12, absTol=0.0)
    S.set_values(vals)
    x, info = S.solve(b)
    assert info.converged and info.finalRelResNorm <= 1e-12
    A = sp.csr_matrix((vals, om.cols, om.rowOffsets))
    xd = spla.spsolve(A.tocsc(), b)
    assert np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-9
    xo, it, rr, _ = O.pcg_amg(om.rowOffsets, om.cols, vals, b, relTol=1e-12, absTol=0.0)
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) < 1e-9
    assert abs(info.iterations - it) <= 2  # same algorithm; dot order differs


def test_pcg_warm_start_zero_rhs_and_mixed(meshes, omeshes):
    m, om = meshes["hex"], omeshes["hex"]
    vals = laplacian(om, m)
    S = lp.PCGSolver(om.rowOffsets, om.cols, relTol=1e-10, absTol=0.0)
    S.set_values(vals)
    x, info = S.solve(np.zeros(om.nCells), x0=np.ones(om.nCells))
    assert np.all(x == 0) and info.iterations == 0  # HYPRE: b == 0 -> x = 0
    b = np.random.default_rng(4).standard_normal(om.nCells)
    x1, i1 = S.solve(b)
    x2, i2 = S.solve(b, x0=x1)  # already converged
    assert i2.iterations == 0 and np.array_equal(x1, x2)
    M = lp.PCGSolver(om.rowOffsets, om.cols, relTol=1e-10, absTol=0.0, precision=1)
    M.set_values(vals)
    xm, im = M.solve(b)
    assert im.converged and im.finalRelResNorm <= 1e-10 and im.refinementSteps >= 2
    assert np.linalg.norm(xm - x1) / np.linalg.norm(x1) < 1e-8
    # precision 2: the FP64 iteration around the FP32 cycle keeps the FP64
    # stop test and about the FP64 iteration count
    P2 = lp.PCGSolver(om.rowOffsets, om.cols, relTol=1e-10, absTol=0.0, precision=2)
    P2.set_values(vals)
    x2p, i2p = P2.solve(b)
    h2 = P2.residual_history()
    assert i2p.converged and i2p.finalRelResNorm <= 1e-10 and i2p.refinementSteps == 0
    assert i2p.iterations <= i1.iterations + 2
    assert len(h2) == i2p.iterations + 1 and h2[-1] == i2p.finalResNorm
    r2 = b - sp.csr_matrix((vals, om.cols, om.rowOffsets)) @ x2p
    assert np.linalg.norm(r2) <= 1.01e-10 * np.linalg.norm(b)
    assert np.linalg.norm(x2p - x1) / np.linalg.norm(x1) < 1e-8


def test_bicgstab_matches_oracle(meshes, omeshes):
    m, om = meshes["cyl"], omeshes["cyl"]
    flux = m.Sf @ np.array([0.5, 0.1, 0.05])  # uniform velocity: conservative fluxes
    t, v = bc_arrays(m, dirichlet=("patch_1_0", "patch_3_0"), value=lambda x: x[:, 1])
    vals, rhs = O.transport_assemble(om, flux, np.full(m.nFaces, 0.01), np.ones(m.nCells),
                                     np.zeros(m.nCells), t, v, np.zeros((m.nCells, 3)))
    x, info = lp.bicgstab(om.rowOffsets, om.cols, vals, rhs, absTol=1e-13, maxIter=5000)
    xo, it, _ = O.bicgstab(om.rowOffsets, om.cols, vals, rhs, absTol=1e-13, maxIter=5000)
    assert info.converged
    A = sp.csr_matrix((vals, om.cols, om.rowOffsets))
    xd = spla.spsolve(A.tocsc(), rhs)
    assert np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-9
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) < 1e-9


def test_bicgstab3_columns_bitwise_equal_single_solves(meshes, omeshes):
    """Fused 3-RHS BiCGSTAB (u/v/w momentum systems sharing one matrix):
    every column equals a separate an_bicgstab_solve bit for bit, including
    its iteration count, when columns converge at different iterations and
    when one right-hand side is zero (x = 0, HYPRE convention)."""
    m, om = meshes["cyl"], omeshes["cyl"]
    flux = m.Sf @ np.array([0.5, 0.1, 0.05])
    t, v = bc_arrays(m, dirichlet=("patch_1_0", "patch_3_0"), value=lambda x: x[:, 1])
    vals, rhs = O.transport_assemble(om, flux, np.full(m.nFaces, 0.01), np.ones(m.nCells),
                                     np.zeros(m.nCells), t, v, np.zeros((m.nCells, 3)))
    rng = np.random.default_rng(12)
    cases = [
        [rhs, 1e-3 * rng.standard_normal(m.nCells), rhs + 0.5 * rng.standard_normal(m.nCells)],
        [rhs, np.zeros(m.nCells), 3.0 * rhs],
    ]
    for bs in cases:
        x0s = [np.zeros(m.nCells), 0.1 * rng.standard_normal(m.nCells), np.ones(m.nCells)]
        xs, infos = lp.bicgstab3(om.rowOffsets, om.cols, vals, bs, x0s, absTol=1e-12,
                                 maxIter=5000)
        its = []
        for c in range(3):
            x1, i1 = lp.bicgstab(om.rowOffsets, om.cols, vals, bs[c], x0s[c], absTol=1e-12,
                                 maxIter=5000)
            assert np.array_equal(xs[c], x1), c
            assert infos[c].iterations == i1.iterations and infos[c].converged == i1.converged
            assert infos[c].finalResNorm == i1.finalResNorm
            its.append(i1.iterations)
        assert infos[0].spmvAvgSeconds > 0.0
    assert len(set(its)) > 1 or 0 in its  # the columns really stopped apart


# --------------------------------------------------- libpoisson drivers
def mms_error(n):
    m = lp.Mesh.tetcube(n)
    bc = lp.ScalarBCSet([lp.ScalarPatchBC(p, L.AN_BC_DIRICHLET,
                                          lambda x, nn: np.prod(np.sin(np.pi * x), axis=1))
                         for p in m.patchNames])
    src = 3 * np.pi ** 2 * np.prod(np.sin(np.pi * m.cc), axis=1)
    phi, info = lp.solve_poisson(m, src, bc, lp.EllipticOptions(nNonOrthCorr=10,
                                                                solver=dict(absTol=1e-12)))
    err = phi - np.prod(np.sin(np.pi * m.cc), axis=1)
    l2 = np.sqrt(np.sum(m.vol * err ** 2) / m.vol.sum())
    return l2, info


def test_poisson_mms_tet_second_order(gpu):
    """apps/poisson_mms_lib (config 1 family): L2 error and observed order."""
    e8, _ = mms_error(8)
    e16, info = mms_error(16)
    assert info.converged
    order = np.log2(e8 / e16)
    assert 1.7 < order < 2.4, (e8, e16, order)


def test_solve_scalar_elliptic_matches_oracle_pipeline(meshes, omeshes):
    m, om = meshes["cyl"], omeshes["cyl"]
    rng = np.random.default_rng(8)
    gam = 1.0 + rng.random(m.nFaces)
    src = rng.standard_normal(m.nCells)
    t, v = bc_arrays(m, dirichlet=("patch_1_0", "patch_2_0"), value=lambda x: x[:, 0] ** 2,
                     neumann_value=lambda x, n: 0.3 * n[:, 1])
    bc = lp.ScalarBCSet()
    for k, name in enumerate(m.patchNames):
        sel = m.bPatch[m.nInternalFaces:] == k + 1
        typ = int(t[sel][0])
        bc.patches.append(lp.ScalarPatchBC(name, typ, (lambda x, n: x[:, 0] ** 2) if typ else
                                           (lambda x, n: 0.3 * n[:, 1])))
    phi, info = lp.solve_scalar_elliptic(m, gam, src, bc, lp.EllipticOptions(
        nNonOrthCorr=3, solver=dict(absTol=1e-13)))
    # oracle pipeline: LP gradient -> assembly -> direct solve, 4 passes from 0
    x = np.zeros(m.nCells)
    for _ in range(4):
        g = O.lsq_gradient(om, x, t, v, "lp")
        vals, rhs = O.elliptic_assemble(om, gam, src, t, v, g, True)
        A = sp.csr_matrix((vals, om.cols, om.rowOffsets))
        x = spla.spsolve(A.tocsc(), rhs)
    assert np.linalg.norm(phi - x) / np.linalg.norm(x) < 1e-8
