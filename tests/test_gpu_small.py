"""Small and ragged systems on the GPU against the oracle.

The other parity tests use meshes of hundreds to thousands of cells.  Here
the fine level itself is smaller than one wavefront or one CSR row block,
the row count is not a multiple of 64, and the AMG hierarchy is either the
dense coarsest level alone (n <= coarseSize) or a chain of levels of a few
rows each (coarseSize 4).  Addressing, gradients and assembly must match the
oracle bit for bit; the hierarchy's application too; PCG follows the
oracle's iteration count and residual history (dot-product order aside;
below 1e-14 of the initial residual only roundoff remains).
"""
import numpy as np
import pytest

from anabasis_amd import _lib as L
from anabasis_amd import libpoisson as lp
from oracle import pyoracle as O

from meshes import bc_arrays

pytestmark = pytest.mark.gpu

# 8, 60, 504 and 1287 cells
SIZES = [(2, 2, 2), (5, 4, 3), (9, 8, 7), (13, 11, 9)]


@pytest.fixture(scope="module", params=SIZES, ids=lambda s: "x".join(map(str, s)))
def box(gpu, request):
    nx, ny, nz = request.param
    m = lp.Mesh.hexbox(nx, ny, nz, 1.0, 0.8, 0.6, jitter=0.2, seed=nx + 7 * ny)
    return m, O.OMesh.from_product(m)


def laplacian(om, m):
    t, v = bc_arrays(m, dirichlet=("xmin",))
    vals, _ = O.elliptic_assemble(om, np.ones(om.nFaces), np.zeros(om.nCells), t, v,
                                  np.zeros((om.nCells, 3)))
    return vals


def test_small_addressing_gradient_assembly_bitexact(box):
    m, om = box
    pat = lp.csr_from_faces(m.nCells, om.owner, om.neigh)
    ref = O.pattern(m.nCells, om.owner, om.neigh)
    for a, b in zip((pat.rowOffsets, pat.cols, pat.diagPos, pat.facePP, pat.facePN, pat.faceNP,
                     pat.faceNN), ref):
        assert np.array_equal(a, b)
    rng = np.random.default_rng(m.nCells)
    nb = m.nBoundaryFaces
    phi = rng.standard_normal(m.nCells)
    t = (rng.random(nb) < 0.5).astype(np.int32)
    v = rng.standard_normal(nb)
    for mode, name in ((L.AN_GRAD_GSV1, "gsv1"), (L.AN_GRAD_LP, "lp")):
        assert np.array_equal(lp.compute_lsq_gradient(m, phi, t, v, mode),
                              O.lsq_gradient(om, phi, t, v, name)), name
    gam = 0.5 + rng.random(m.nFaces)
    grad = rng.standard_normal((m.nCells, 3))
    vals, rhs = lp.assemble_scalar_elliptic_system(m, gam, phi, t, v, grad, True, False, 3, 1.25,
                                                   nnz=om.nnz)
    rv, rr = O.elliptic_assemble(om, gam, phi, t, v, grad, True, False, 3, 1.25)
    assert np.array_equal(vals, rv) and np.array_equal(rhs, rr)


@pytest.mark.parametrize("coarse", [64, 4])
def test_small_spmv_and_hierarchy_bitexact(box, coarse):
    m, om = box
    vals = laplacian(om, m)
    S = lp.PCGSolver(om.rowOffsets, om.cols, coarseSize=coarse, kcycle=0)
    S.set_values(vals)
    x = np.random.default_rng(3).standard_normal(om.nCells)
    assert np.array_equal(S.spmv(x), O.spmv(om.rowOffsets, om.cols, vals, x))
    z = S.amg_apply(x)
    zo, nl = O.amg_vcycle(om.rowOffsets, om.cols, vals, x, coarseSize=coarse, kcycle=0)
    if om.nCells <= coarse:
        assert nl == 1  # the dense coarsest level alone
    assert np.array_equal(z, zo)


@pytest.mark.parametrize("coarse", [64, 4])
def test_small_pcg_follows_oracle(box, coarse):
    m, om = box
    vals = laplacian(om, m)
    b = np.random.default_rng(17).standard_normal(om.nCells)
    S = lp.PCGSolver(om.rowOffsets, om.cols, relTol=1e-10, absTol=0.0, coarseSize=coarse)
    S.set_values(vals)
    x, info = S.solve(b)
    h = S.residual_history()
    xo, it, rr, ho = O.pcg_amg(om.rowOffsets, om.cols, vals, b, relTol=1e-10, absTol=0.0,
                               coarseSize=coarse, hist_cap=500)
    assert info.converged and info.iterations == it and len(h) == it + 1
    ho = ho[: it + 1]
    # an exact coarsest solve (one level) or an exhausted Krylov space ends
    # at roundoff: there the floor 1e-14 ||r0|| applies, elsewhere 1e-8
    # relative
    assert np.all(np.abs(h - ho) <= 1e-8 * ho + 1e-14 * ho[0]), np.abs(h - ho) / ho
    assert np.linalg.norm(x - xo) <= 1e-9 * np.linalg.norm(xo)


@pytest.mark.parametrize("coarse", [64, 4])
def test_small_mixed_precision_pcg(box, coarse):
    """FP32 inner PCG/AMG with FP64 refinement on the same small systems
    (the FP32 copy of a dense-only or few-row hierarchy): converged to the
    FP64 tolerance and at the oracle's FP64 solution."""
    m, om = box
    vals = laplacian(om, m)
    b = np.random.default_rng(17).standard_normal(om.nCells)
    S = lp.PCGSolver(om.rowOffsets, om.cols, relTol=1e-10, absTol=0.0, coarseSize=coarse,
                     precision=1)
    S.set_values(vals)
    x, info = S.solve(b)
    xo, it, rr, _ = O.pcg_amg(om.rowOffsets, om.cols, vals, b, relTol=1e-10, absTol=0.0,
                              coarseSize=coarse)
    assert info.converged and info.finalRelResNorm <= 1e-10
    assert np.linalg.norm(x - xo) <= 1e-8 * np.linalg.norm(xo)


def test_small_bicgstab_and_fused_columns(box):
    """Momentum-type (non-symmetric) systems: BiCGSTAB reaches the oracle's
    solution, and the fused 3-RHS solve equals three single solves bit for
    bit (the element kernels take elements in pairs, so odd and sub-block
    row counts exercise their tails).  A small non-symmetric system can break
    down (the 60-cell box's third column does, in the oracle too, after 78
    iterations): the fused solve then reports that column unconverged exactly
    as the single solve does."""
    m, om = box
    flux = m.Sf @ np.array([0.5, 0.1, 0.05])
    t, v = bc_arrays(m, dirichlet=("xmin", "xmax"), value=lambda x: x[:, 1])
    vals, rhs = O.transport_assemble(om, flux, np.full(m.nFaces, 0.01), np.ones(m.nCells),
                                     np.zeros(m.nCells), t, v, np.zeros((m.nCells, 3)))
    x, info = lp.bicgstab(om.rowOffsets, om.cols, vals, rhs, absTol=1e-13, maxIter=5000)
    xo, it, _ = O.bicgstab(om.rowOffsets, om.cols, vals, rhs, absTol=1e-13, maxIter=5000)
    assert info.converged
    assert np.linalg.norm(x - xo) <= 1e-9 * np.linalg.norm(xo)
    rng = np.random.default_rng(m.nCells + 1)
    bs = [rhs, np.zeros(m.nCells), rhs + 0.5 * rng.standard_normal(m.nCells)]
    x0s = [np.zeros(m.nCells), 0.1 * rng.standard_normal(m.nCells), np.ones(m.nCells)]
    xs, infos = lp.bicgstab3(om.rowOffsets, om.cols, vals, bs, x0s, absTol=1e-12, maxIter=5000,
                             allow_noconv=True)
    for c in range(3):
        x1, i1 = lp.bicgstab(om.rowOffsets, om.cols, vals, bs[c], x0s[c], absTol=1e-12,
                             maxIter=5000, allow_noconv=True)
        assert np.array_equal(xs[c], x1), c
        assert infos[c].iterations == i1.iterations and infos[c].converged == i1.converged
        xo, ito, _ = O.bicgstab(om.rowOffsets, om.cols, vals, bs[c], x0s[c], absTol=1e-12,
                                maxIter=5000)
        if c < 2:
            assert infos[c].converged, c
        if infos[c].converged and ito >= 0:
            assert np.linalg.norm(xs[c] - xo) <= 1e-9 * max(np.linalg.norm(xo), 1e-300), c


@pytest.mark.parametrize("dims", [(5, 4, 3), (9, 8, 7)], ids=["60", "504"])
def test_small_simple_matches_oracle(gpu, dims):
    """Four SIMPLE outer iterations of a small box channel (uniform inflow at
    xmin, fixed pressure at xmax, no-slip elsewhere; a single-level or
    two-level pressure hierarchy) against the oracle: fields within 1e-8,
    wall forces on the same fields bit for bit."""
    from anabasis_amd import simple as sm
    m = lp.Mesh.hexbox(*dims, 1.0, 0.8, 0.6, jitter=0.2, seed=dims[0])
    vel = [sm.VelocityPatchBC("xmin", "fixed_uniform_vector", vector=(0.3, 0.0, 0.0)),
           sm.VelocityPatchBC("xmax", "zero_gradient")]
    vel += [sm.VelocityPatchBC(p, "wall_noslip") for p in ("ymin", "ymax", "zmin", "zmax")]
    prs = [sm.PressurePatchBC(p, "zero_gradient") for p in ("xmin", "ymin", "ymax", "zmin",
                                                             "zmax")]
    prs.append(sm.PressurePatchBC("xmax", "fixed_value", 0.0))
    # mu = 0.01 keeps the cell Reynolds number of these coarse boxes low
    # enough for the iteration to converge
    par = sm.default_params(nNonOrthCorr=2, velTol=1e-14, pTol=1e-14, mu=0.01,
                            pressure=dict(coarseSize=64))
    S = sm.Simple(m, par, vel, prs)
    om = O.OMesh.from_product(m)
    tU, tP, ub, vb, wb, pb = S.bc
    u0 = [S.field(k) for k in "uvwp"]
    st, _ = S.iterate(4)
    ref = O.simple_run(om, O.simple_params_from(par), tU, tP, ub, vb, wb, pb, 4, *u0)
    assert ref["done"] == 4 and st.step == 4
    for k in ("u", "v", "w", "p", "phi"):
        a, b = S.field(k), ref[k]
        assert np.linalg.norm(a - b) <= 1e-8 * np.linalg.norm(b), k
    assert np.array_equal(S.field("pA"), O.pressure_assemble(om, S.field("rAU"), tP))
    k = m.patch_index("ymin")
    f = S.patch_forces("ymin", -1)
    fo = O.patch_forces(om, m.patchStartFace[k], m.patchStartFace[k] + m.patchNFaces[k],
                        S.field("u"), S.field("v"), S.field("w"), S.field("p"), par.mu, -1)
    assert np.array_equal(f, fo)
