"""The libpoisson / libscalar C++ headers over the C ABI: they compile with
the reference's own call pattern (a port of apps/poisson_mms_lib/main.cu)
and, on the GPU, that driver reproduces the MMS results of the Python API
with second-order convergence on tet meshes."""
import os
import pathlib
import re
import shutil
import subprocess

import numpy as np
import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
LIBDIR = ROOT / "anabasis_amd"
SRC = ROOT / "tests" / "cpp" / "poisson_mms_lib.cpp"


def build_driver(out_dir):
    exe = pathlib.Path(out_dir) / "poisson_mms_lib"
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-Werror", f"-I{ROOT / 'include'}",
           str(SRC), "-o", str(exe), f"-L{LIBDIR}", "-lanabasis_hip", f"-Wl,-rpath,{LIBDIR}"]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return exe


def test_headers_compile_and_link(tmp_path):
    if not (LIBDIR / "libanabasis_hip.so").exists():
        pytest.fail("libanabasis_hip.so is not built")
    exe = build_driver(tmp_path)
    assert exe.exists()
    # libscalar header: the reference's call pattern (scalar_transport_library.h)
    tu = tmp_path / "ls.cpp"
    tu.write_text(
        "#include <anabasis/libscalar.hpp>\n"
        "int run(const Mesh& m) {\n"
        "  libscalar::ScalarBCSet bc;\n"
        "  bc.patches.push_back(libscalar::make_dirichlet_constant_bc(\"xmin\", 1.0));\n"
        "  bc.patches.push_back(libscalar::make_zero_gradient_patch_bc(\"xmax\"));\n"
        "  libscalar::ScalarTransportInputs in;\n"
        "  libscalar::ScalarTransportOptions opt;\n"
        "  opt.convectionScheme = libscalar::ConvectionScheme::Upwind;\n"
        "  auto r = libscalar::solve_steady_scalar_transport(m, in, bc, opt);\n"
        "  return r.iterations;\n"
        "}\n")
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-fsyntax-only",
                    f"-I{ROOT / 'include'}", str(tu)], check=True)


def run_driver(exe, mesh_dir, extra=()):
    r = subprocess.run([str(exe), "-polyMeshDir", str(mesh_dir), "-tol", "1e-12", *extra],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = {}
    for key in ("L2 error", "Linf error", "last hypreIts", "last relRes"):
        m = re.search(rf"{key}\s*:\s*(\S+)", r.stdout)
        out[key] = float(m.group(1))
    return out, r.stdout


@pytest.mark.gpu
def test_cpp_mms_driver_matches_python_api(gpu, tmp_path):
    from anabasis_amd import libpoisson as lp
    exe = build_driver(tmp_path)
    errs = []
    for n in (8, 16, 32):
        m = lp.Mesh.tetcube(n)
        d = tmp_path / f"tet{n}"
        m.write_polymesh(d)
        res, text = run_driver(exe, d)
        assert f"cells={6 * n ** 3}" in text
        errs.append(res["L2 error"])
        if n == 8:
            # same pipeline through the Python API (ctypes), same defaults
            u = lambda x: np.sin(np.pi * x[:, 0]) * np.sin(np.pi * x[:, 1]) * np.sin(np.pi * x[:, 2])
            bcs = lp.ScalarBCSet([lp.ScalarPatchBC(p, lp.AN_BC_DIRICHLET, lambda x, nrm: u(x))
                                  for p in m.patchNames])
            cc, vol = m.array("cc"), m.array("vol")
            opts = lp.EllipticOptions(nNonOrthCorr=10, solver=dict(absTol=1e-12, relTol=0.0))
            phi, _ = lp.solve_scalar_elliptic(m, np.ones(m.nFaces), 3 * np.pi ** 2 * u(cc), bcs,
                                              opts)
            l2 = np.sqrt(np.sum(vol * (phi - u(cc)) ** 2) / vol.sum())
            assert abs(l2 - res["L2 error"]) <= 1e-9 * l2
    orders = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert np.all(orders > 1.8), (errs, orders)
