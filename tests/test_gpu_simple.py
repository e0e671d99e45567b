"""SIMPLE outer iterations on the GPU against the oracle's restatement of
generic_simple_v1 (main.cu:3388-3634) on the cylinder-channel geometry."""
import numpy as np
import pytest

from anabasis_amd import libpoisson as lp
from anabasis_amd import simple as sm
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.fixture(scope="module")
def case(gpu):
    m = lp.Mesh.cylinder(6000)
    vel, prs = sm.cylinder_case_bcs()
    sm.coverage_table(m, vel, prs)
    # tight solver tolerances so both sides reach the same converged fields
    par = sm.default_params(nNonOrthCorr=2, velTol=1e-14, pTol=1e-14,
                            pressure=dict(coarseSize=64))
    return m, vel, prs, par


def test_simple_matches_oracle(case):
    m, vel, prs, par = case
    S = sm.Simple(m, par, vel, prs)
    om = O.OMesh.from_product(m)
    tU, tP, ub, vb, wb, pb = S.bc
    u0 = [S.field(k) for k in "uvwp"]
    n = 4
    st, conv = S.iterate(n)
    ref = O.simple_run(om, O.simple_params_from(par), tU, tP, ub, vb, wb, pb, n, *u0)
    assert ref["done"] == n and st.step == n
    for k in "uvwp":
        assert rel(S.field(k), ref[k]) < 1e-8, k
    assert rel(S.field("phi"), ref["phi"]) < 1e-8
    assert rel(S.field("rAU"), ref["rAU"]) < 1e-10
    # the last pressure matrix equals the oracle's assembly from the same
    # rAU bit for bit (k_p_face + k_p_rows against GSV1 main.cu:1898-1950)
    assert np.array_equal(S.field("pA"), O.pressure_assemble(om, S.field("rAU"), tP))
    rs = ref["stats"][-1]
    assert abs(st.massRes - rs.massRes) <= 1e-6 * max(rs.massRes, 1e-30) + 1e-12
    # forces on the cylinder (forceNormalSign -1, cases/cylinder.case:68-70)
    k = m.patch_index("patch_3_0")
    f = S.patch_forces("patch_3_0", -1)
    fo = O.patch_forces(om, m.patchStartFace[k], m.patchStartFace[k] + m.patchNFaces[k],
                        S.field("u"), S.field("v"), S.field("w"), S.field("p"), par.mu, -1)
    assert np.array_equal(f, fo)  # same fields -> identical serial sums


def test_simple_converged_state_matches_oracle(gpu):
    """Run to the case's own stop test (massRes < tolMass and the largest
    velocity change < tolVel) on both sides: the same iteration count, the
    converged fields within 1e-8 and the cylinder force within 1e-8."""
    m = lp.Mesh.cylinder(20000)
    vel, prs = sm.cylinder_case_bcs()
    par = sm.default_params(velTol=1e-13, pTol=1e-13, pressure=dict(coarseSize=64))
    S = sm.Simple(m, par, vel, prs)
    om = O.OMesh.from_product(m)
    tU, tP, ub, vb, wb, pb = S.bc
    u0 = [S.field(k) for k in "uvwp"]
    st, conv = S.iterate(300)
    ref = O.simple_run(om, O.simple_params_from(par), tU, tP, ub, vb, wb, pb, 300, *u0)
    assert conv and st.step == ref["done"] and st.step < 300
    for k in "uvwp":
        assert rel(S.field(k), ref[k]) < 1e-8, k
    k = m.patch_index("patch_3_0")
    f = S.patch_forces("patch_3_0", -1)
    fo = O.patch_forces(om, m.patchStartFace[k], m.patchStartFace[k] + m.patchNFaces[k],
                        ref["u"], ref["v"], ref["w"], ref["p"], par.mu, -1)
    fc, foc = np.asarray(f)[6:9], np.asarray(fo)[6:9]  # total force on the cylinder
    assert np.abs(fc - foc).max() <= 1e-8 * np.abs(foc).max()


def test_simple_flux_is_conservative(case):
    m, vel, prs, par = case
    S = sm.Simple(m, par, vel, prs)
    st, _ = S.iterate(2)
    phi = S.field("phi")
    div = np.zeros(m.nCells)
    nIF = m.nInternalFaces
    np.add.at(div, m.owner, phi)
    np.add.at(div, m.neigh, -phi[:nIF])
    assert np.abs(div).max() == pytest.approx(st.massRes, rel=1e-12, abs=1e-15)
    # what remains is the change of the explicit non-orthogonal flux between
    # the last solve and the final correction (reference behaviour)
    assert st.massRes < 1e-4


def test_simple_inert_without_driving(case):
    """rho = 0, zero inlet velocity: nothing moves, pressure stays zero."""
    m, _, prs, par0 = case
    vel = [sm.VelocityPatchBC(n, "wall_noslip") for n in m.patchNames if n != "patch_2_0"]
    vel.append(sm.VelocityPatchBC("patch_2_0", "zero_gradient"))
    S = sm.Simple(m, sm.default_params(nNonOrthCorr=1), vel, prs)
    st, _ = S.iterate(2)
    for k in "uvwp":
        assert np.abs(S.field(k)).max() < 1e-12


def test_simple_fused_momentum_is_bitwise_identical(case):
    """fuseMomentumRhs: u/v/w advance in lock-step through one 3-RHS
    BiCGSTAB per pass; every field equals the component-by-component
    iteration bit for bit."""
    m, vel, prs, par = case
    out = []
    for fuse in (0, 1):
        p2 = sm.default_params(nNonOrthCorr=2, velTol=1e-14, pTol=1e-14, fuseMomentumRhs=fuse,
                               pressure=dict(coarseSize=64))
        S = sm.Simple(m, p2, vel, prs)
        st, _ = S.iterate(3)
        out.append(({k: S.field(k) for k in ("u", "v", "w", "p", "phi", "rAU")},
                    list(st.bicgstabIts), st.massRes))
    for k in out[0][0]:
        assert np.array_equal(out[0][0][k], out[1][0][k]), k
    assert out[0][1] == out[1][1] and out[0][2] == out[1][2]


@pytest.mark.parametrize("precision", [1, 2])
def test_mixed_precision_converged_state_matches_fp64_oracle(gpu, precision):
    """The mixed-precision pressure solves -- 1: FP32 PCG + AMG inside FP64
    refinement; 2: FP64 PCG preconditioned by the FP32 AMG cycle -- run to
    the case's own stop test on the 16,660-cell cylinder channel: the same
    stop iteration as the FP64 oracle, fields within 1e-9 relative and the
    cylinder force within 1e-8 of |F| (measured, precision 1: fields 4e-11,
    force 9.4e-12 |F|)."""
    m = lp.Mesh.cylinder(20000)
    vel, prs = sm.cylinder_case_bcs()
    par = sm.default_params(velTol=1e-13, pTol=1e-13, precision=precision,
                            pressure=dict(coarseSize=64))
    S = sm.Simple(m, par, vel, prs)
    om = O.OMesh.from_product(m)
    u0 = [S.field(k) for k in "uvwp"]
    st, conv = S.iterate(300)
    par64 = sm.default_params(velTol=1e-13, pTol=1e-13, pressure=dict(coarseSize=64))
    ref = O.simple_run(om, O.simple_params_from(par64), *S.bc, 300, *u0)
    assert conv and st.step == ref["done"] < 300
    for k in "uvwp":
        assert rel(S.field(k), ref[k]) < 1e-9, k
    k = m.patch_index("patch_3_0")
    f = np.asarray(S.patch_forces("patch_3_0", -1))[6:9]
    fo = np.asarray(O.patch_forces(om, m.patchStartFace[k], m.patchStartFace[k] + m.patchNFaces[k],
                                   ref["u"], ref["v"], ref["w"], ref["p"], par.mu, -1))[6:9]
    assert np.abs(f - fo).max() <= 1e-8 * np.linalg.norm(fo)
