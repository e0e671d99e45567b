"""The generic_simple_v1 command line (anabasis_amd/generic_simple_v1):

CPU (all of it happens before the GPU is touched): the .case grammar and its
error texts (GSV1 main.cu:162-331), command-line flags overriding case keys,
the generated BC file, the missing-BC and missing-patch errors (main.cu:
2962-2990, 3108-3119, 3178-3233) and the flag errors (main.cu:111-124,
389-397).

GPU: a run of the cylinder channel: the iteration log line format
(main.cu:3608-3612), the stop message, the final summary and the
_summary.txt keys (main.cu:3844-3904) with the same values as the same run
through the Python API, the paper-formula cylinder force against an
independent numpy evaluation of its formula (main.cu:864-941), and a
well-formed polyhedron VTU file (main.cu:975-983)."""
import re
import subprocess
import xml.etree.ElementTree as ET

import numpy as np
import pytest

from conftest import ROOT

CLI = ROOT / "anabasis_amd" / "generic_simple_v1"

CASE = """# generated cylinder channel
polyMeshDir {mesh}
outPrefix {prefix}
device 0
rho 1.0
mu 0.001
Umean 0.45
pipeD 0.1
pipeL 2.5
wallPatch patch_0_0
inletPatch patch_1_0
outletPatch patch_2_0
nsteps {nsteps}
printEvery 5
writeVtu 1
writeEvery 0
monitor 1
nVelNonOrthCorr 1
nNonOrthCorr 2
nPressureCorr 0
uRelax 0.5
pRelax 0.3
corrTol 1e-10
tolMass 1e-6
tolVel 1e-6
velMaxit 1000
velTol 1e-12
velRelTol 0
pUseAmg 1
pMaxit 4000
pTol 1e-12
pRelTol 0
pAmgRebuildEvery 10
{bc}
forceEnable 1
forcePatch patch_3_0
forceNormalSign -1
"""

BCS = """velocity patch_0_0 wall_noslip
velocity patch_5_0 wall_noslip
velocity patch_3_0 wall_noslip
velocity patch_1_0 parabolic_box_inlet 0.41 0.45 y z average_patch_normal
velocity patch_2_0 zero_gradient
pressure patch_0_0 zero_gradient
pressure patch_5_0 zero_gradient
pressure patch_3_0 zero_gradient
pressure patch_1_0 zero_gradient
pressure patch_2_0 fixed_value 0.0"""


@pytest.fixture(scope="module")
def cyl_mesh(tmp_path_factory):
    from anabasis_amd import libpoisson as lp
    if not CLI.exists():
        pytest.fail(f"{CLI} is not built")
    d = tmp_path_factory.mktemp("cyl") / "polyMesh"
    m = lp.Mesh.cylinder(6000)
    m.write_polymesh(d)
    return m, d


def write_case(tmp_path, mesh_dir, bc=BCS, nsteps=40, extra=""):
    case = tmp_path / "cyl.case"
    case.write_text(CASE.format(mesh=mesh_dir, prefix=tmp_path / "out", nsteps=nsteps, bc=bc) +
                    extra)
    return case


def run(*args, timeout=120):
    return subprocess.run([str(CLI), *map(str, args)], capture_output=True, text=True,
                          timeout=timeout)


@pytest.mark.parametrize("text,msg", [
    ("rho 1.0 2.0\n", "at line 1: expected '<key> <value>' or a velocity/pressure BC line"),
    ("nsteps 5\ndensity 1.0\n", "at line 2: unknown key 'density'"),
    ("bcConfig a.bc\nvelocity p wall_noslip\n",
     "Case config cannot contain both bcConfig and inline velocity/pressure BC lines"),
])
def test_case_grammar_errors(tmp_path, text, msg):
    case = tmp_path / "bad.case"
    case.write_text(text)
    r = run("-case-config", case)
    assert r.returncode == 1
    if "at line" in msg:
        msg = f"Case config parse error in '{case}' " + msg
    assert r.stderr.strip() == msg


def test_flag_errors(tmp_path):
    r = run("-case-config")
    assert r.returncode == 1 and r.stderr.strip() == "Missing value after -case-config"
    r = run("-case-config", tmp_path / "none.case")
    assert r.stderr.strip() == f"Could not open case config file: {tmp_path / 'none.case'}"
    r = run("-p-amg-setup-scope", "everywhere")
    assert r.stderr.strip() == "Unknown -p-amg-setup-scope 'everywhere'. Use outer or pressure."
    r = run("-force-drag-dir", "1,0")
    assert r.stderr.strip() == "Could not parse -force-drag-dir value '1,0'. Expected x,y,z"
    r = run("-nsteps")
    assert r.stderr.strip() == "Missing value after -nsteps"


def test_case_expansion_and_input_errors(tmp_path, cyl_mesh):
    m, mesh_dir = cyl_mesh
    # inline BC lines -> <case>.generated.bc
    case = write_case(tmp_path, mesh_dir, extra="nPressureCorr 2\n")
    r = run("-case-config", case)
    assert r.returncode == 1 and "nPressureCorr > 0" in r.stderr
    gen = tmp_path / "cyl.case.generated.bc"
    assert gen.read_text() == f"# Auto-generated from {case}\n" + BCS + "\n"
    # command-line flags override the case file (they come after its keys)
    case = write_case(tmp_path, mesh_dir, bc=BCS.replace("pressure patch_3_0 zero_gradient\n", "")
                      .replace("velocity patch_2_0 zero_gradient\n", ""))
    r = run("-case-config", case, "-nsteps", 7, "-mu", 0.002)
    assert "nsteps         : 7\n" in r.stdout
    assert "mu             : 2.00000000e-03\n" in r.stdout
    assert "mu source      : direct -mu / case mu\n" in r.stdout
    assert f"Cells          : {m.nCells}\n" in r.stdout
    # incomplete coverage: the table on stdout, the report on stderr, before the GPU
    assert r.returncode == 1
    assert "  patch_2_0                        velocity=NO pressure=yes\n" in r.stdout
    assert "  patch_3_0                        velocity=yes pressure=NO\n" in r.stdout
    assert r.stderr == ("\nERROR: incomplete runtime BC coverage.\n"
                        "Missing velocity BC for patches:\n  patch_2_0\n"
                        "Missing pressure BC for patches:\n  patch_3_0\n"
                        "\nEvery polyMesh boundary patch must have exactly one velocity BC and "
                        "one pressure BC in generic_simple_v1.\n")
    assert "Running on HIP device" not in r.stdout
    # missing wall/inlet/outlet patch, missing BCs
    r = run("-case-config", case, "-wall-patch", "walls")
    assert r.returncode == 1 and r.stderr.strip() == "Could not find wall/inlet/outlet patch."
    r = run("-polyMeshDir", mesh_dir)
    assert r.returncode == 1 and r.stderr.strip() == (
        "generic_simple_v1 requires explicit runtime BCs.\n"
        "Add velocity/pressure lines to the case file, or set bcConfig.")
    # a BC grammar error comes back with the reference's text
    bad = tmp_path / "bad.bc"
    bad.write_text("velocity patch_1_0 parabolic_box_inlet 0.41 0.45 y y\n")
    r = run("-polyMeshDir", mesh_dir, "-bc-config", bad)
    assert r.returncode == 1
    assert r.stderr.strip().endswith(f"BC config parse error in '{bad}' at line 1: "
                                     "parabolic_box_inlet requires two different transverse "
                                     "coordinates")


LOG = re.compile(r"^iter +(\d+) : massRes = (\S+), duRel = \S+, dvRel = \S+, dwRel = \S+, "
                 r"dpRel = \S+, bicgstab=\[\d+ \d+ \d+\], pcgLast=\d+, pcgTot=\d+, "
                 r"iterWall = \S+ s, totalWall = \S+ s$")


@pytest.mark.gpu
def test_cli_run_matches_python_api(gpu, tmp_path, cyl_mesh):
    from anabasis_amd import simple as sm
    m, mesh_dir = cyl_mesh
    case = write_case(tmp_path, mesh_dir, nsteps=12)
    r = run("-case-config", case, timeout=300)
    assert r.returncode == 0, r.stderr
    out = r.stdout
    steps = [int(g.group(1)) for g in (LOG.match(l) for l in out.splitlines()) if g]
    assert steps == [1, 5, 10]  # step 1 and every printEvery
    assert "Runtime BC patch coverage:\n" in out and "Pressure reference required: no\n" in out
    summ = {}
    for line in (tmp_path / "out_summary.txt").read_text().splitlines():
        k, v = line.split(" ", 1)
        summ[k] = v
    for k in ("points", "faces", "internalFaces", "cells", "maxNonOrthDeg", "mu", "steps",
              "massRes", "maxAbsU", "maxAbsV", "maxAbsW", "maxAbsP", "lastBiCGSTABU",
              "lastBiCGSTABV", "lastBiCGSTABW", "lastPCG", "cylinderForcePatch",
              "cylinderForceFaces", "cylinderForceArea", "forceUbar", "forceD", "forceH",
              "forceCoeffDenom", "FDPressure", "FDViscous", "FDTotal", "FLPressure", "FLViscous",
              "FLTotal", "CD", "CL", "wallDnMin", "wallDnMax", "maxAbsDvtDn", "CDVector",
              "CLyVector", "CLzVector", "FxPressureVector", "FyPressureVector",
              "FzPressureVector", "FxViscousVector", "FyViscousVector", "FzViscousVector",
              "FxTotalVector", "FyTotalVector", "FzTotalVector", "patchForcePatch",
              "patchForceFaces", "patchForceArea", "patchForceNormalSign", "pressureForce_x",
              "pressureForce_y", "pressureForce_z", "viscousForce_x", "viscousForce_y",
              "viscousForce_z", "totalForce_x", "totalForce_y", "totalForce_z"):
        assert k in summ, k
    assert int(summ["cells"]) == m.nCells and int(summ["steps"]) == 12
    # the same run through the Python API: the same library, the same bits
    vel, prs = sm.cylinder_case_bcs()
    par = sm.default_params(rho=1.0, mu=1e-3, nVelNonOrthCorr=1, nNonOrthCorr=2, uRelax=0.5,
                            pRelax=0.3, corrTol=1e-10, tolMass=1e-6, tolVel=1e-6, velMaxit=1000,
                            velTol=1e-12, velRelTol=0.0, pMaxit=4000, pTol=1e-12, pRelTol=0.0,
                            pAmgRebuildEvery=10)
    S = sm.Simple(m, par, vel, prs)
    st, _ = S.iterate(12)
    F = S.patch_forces("patch_3_0", -1)
    assert float(summ["FxTotalVector"]) == pytest.approx(F[6], rel=1e-15, abs=0)
    assert float(summ["totalForce_y"]) == pytest.approx(F[7], rel=1e-15, abs=0)
    assert float(summ["massRes"]) == pytest.approx(st.massRes, rel=1e-15)
    u, v, w, p = (S.field(k) for k in "uvwp")
    assert float(summ["maxAbsP"]) == pytest.approx(np.abs(p).max(), rel=1e-15)
    # C = 2 F / (rho Ubar^2 D H), Ubar = 4/9 Umean, H = 0.41
    Ubar = 4.0 / 9.0 * 0.45
    assert float(summ["CDVector"]) == pytest.approx(2 * F[6] / (Ubar ** 2 * 0.1 * 0.41),
                                                    rel=1e-14)
    # paper formula, evaluated independently from the fields
    k = m.patch_index("patch_3_0")
    f = np.arange(m.patchStartFace[k], m.patchStartFace[k] + m.patchNFaces[k])
    P = m.owner[f]
    n = -m.array("nf")[f]
    n /= np.linalg.norm(n, axis=1)[:, None]
    t = np.stack([n[:, 1], -n[:, 0], 0 * n[:, 0]], 1)
    t /= np.linalg.norm(t, axis=1)[:, None]
    dx = m.array("cc")[P] - m.array("xf")[f]
    dn = np.sum(dx * n, 1)
    dn = np.where(dn <= 1e-14, np.abs(dn), dn)
    vt = u[P] * t[:, 0] + v[P] * t[:, 1] + w[P] * t[:, 2]
    A = m.array("Af")[f]
    FD = np.sum(-p[P] * n[:, 0] * A) + np.sum(1e-3 * vt / dn * n[:, 1] * A)
    assert float(summ["FDTotal"]) == pytest.approx(FD, rel=1e-12)
    # the VTU: one polyhedron per cell with consistent offsets and cell data
    root = ET.parse(tmp_path / "out_final.vtu").getroot()
    piece = root.find("UnstructuredGrid/Piece")
    assert int(piece.get("NumberOfCells")) == m.nCells
    arrays = {a.get("Name"): a.text.split() for a in piece.iter("DataArray") if a.get("Name")}
    assert arrays["offsets"][-1] == str(len(arrays["connectivity"]))
    assert arrays["faceoffsets"][-1] == str(len(arrays["faces"]))
    assert set(arrays["types"]) == {"42"}
    pv = np.array(arrays["p"], float)
    assert np.allclose(pv, p, rtol=1e-14, atol=1e-300)
    assert len(arrays["U"]) == 3 * m.nCells and len(arrays["cell_volume"]) == m.nCells
