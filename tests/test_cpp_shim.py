"""The libpoisson / libscalar C++ headers over the C ABI, driven by
tests/cpp/shim_mms.cpp the way the reference's library callers drive them:

* Dirichlet Poisson MMS (apps/poisson_mms_lib/main.cu:115-214),
* mixed Dirichlet/Neumann MMS (apps/poisson_mms_mixed_lib/main.cu:150-254),
* pure-Neumann MMS with source-mean removal and a reference cell
  (apps/poisson_mms_neumann_lib/main.cu:157-200),
* libscalar::solve_steady_scalar_transport on the scalar_mms_lib case
  (apps/scalar_mms_lib/main.cu:16-30, 86-125: U = (1+y, 0.5-x, 0.25),
  gamma = 0.1, central and upwind convection).

On the CPU the driver must compile and link with -Werror.  On the GPU each
case must reproduce the Python API's result on the same mesh (the same
library underneath, reached through a different host layer) and show the
convergence order the discretisation has: 2 for central differences and the
elliptic cases, 1 for upwind."""
import pathlib
import subprocess

import numpy as np
import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
LIBDIR = ROOT / "anabasis_amd"
SRC = ROOT / "tests" / "cpp" / "shim_mms.cpp"


def build_driver(out_dir):
    exe = pathlib.Path(out_dir) / "shim_mms"
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-Werror", f"-I{ROOT / 'include'}",
           str(SRC), "-o", str(exe), f"-L{LIBDIR}", "-lanabasis_hip", f"-Wl,-rpath,{LIBDIR}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def run_driver(exe, *args):
    r = subprocess.run([str(exe), *map(str, args)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = {}
    for line in r.stdout.splitlines():
        k, v = line.split()
        out[k] = float(v)
    return out


def test_headers_compile_and_link(tmp_path):
    if not (LIBDIR / "libanabasis_hip.so").exists():
        pytest.fail("libanabasis_hip.so is not built")
    assert build_driver(tmp_path).exists()


def test_driver_rejects_bad_arguments(tmp_path):
    """Argument errors are reported before any device work (no GPU needed)."""
    exe = build_driver(tmp_path)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr
    r = subprocess.run([str(exe), "poisson", str(tmp_path / "missing")], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 1 and "error:" in r.stderr


def _orders(errs):
    e = np.asarray(errs)
    return np.log2(e[:-1] / e[1:])


def _sin3(x):
    return np.sin(np.pi * x[:, 0]) * np.sin(np.pi * x[:, 1]) * np.sin(np.pi * x[:, 2])


def _cos3(x):
    return np.cos(np.pi * x[:, 0]) * np.cos(np.pi * x[:, 1]) * np.cos(np.pi * x[:, 2])


def _l2(mesh, phi, exact):
    vol = mesh.array("vol")
    return float(np.sqrt(np.sum(vol * (phi - exact(mesh.array("cc"))) ** 2) / vol.sum()))


@pytest.mark.gpu
def test_cpp_poisson_mms_matches_python_api(gpu, tmp_path):
    from anabasis_amd import libpoisson as lp
    exe = build_driver(tmp_path)
    errs = []
    for n in (8, 16, 32):
        m = lp.Mesh.tetcube(n)
        d = tmp_path / f"tet{n}"
        m.write_polymesh(d)
        res = run_driver(exe, "poisson", d)
        assert res["cells"] == 6 * n ** 3
        errs.append(res["l2"])
        if n == 8:
            bcs = lp.ScalarBCSet([lp.ScalarPatchBC(p, lp.AN_BC_DIRICHLET,
                                                   lambda x, nrm: _sin3(x))
                                  for p in m.patchNames])
            opts = lp.EllipticOptions(nNonOrthCorr=10, solver=dict(absTol=1e-12, relTol=0.0))
            phi, _ = lp.solve_poisson(m, 3 * np.pi ** 2 * _sin3(m.array("cc")), bcs, opts)
            l2 = _l2(m, phi, _sin3)
            assert abs(l2 - res["l2"]) <= 1e-9 * l2
    assert np.all(_orders(errs) > 1.8), errs


@pytest.mark.gpu
def test_cpp_mixed_and_neumann_mms(gpu, tmp_path):
    """cos(pi x)cos(pi y)cos(pi z) on the unit cube; its normal derivative
    vanishes on every face, so Neumann data are zero and LP's Neumann sign
    (rhs -= gamma g A, LP/src/scalar_elliptic.cu:121-123) does not enter."""
    from anabasis_amd import libpoisson as lp
    exe = build_driver(tmp_path)
    mixed, neu = [], []
    for n in (8, 16, 32):
        m = lp.Mesh.tetcube(n)
        d = tmp_path / f"tet{n}"
        m.write_polymesh(d)
        r1 = run_driver(exe, "mixed", d, "xmin,ymax")
        r2 = run_driver(exe, "neumann", d, 0)
        mixed.append(r1["l2"])
        neu.append(r2["l2"])
        if n == 8:
            cc, vol = m.array("cc"), m.array("vol")
            f = 3 * np.pi ** 2 * _cos3(cc)
            # the same mixed problem through the Python API
            bcs = lp.ScalarBCSet([lp.ScalarPatchBC(p, lp.AN_BC_DIRICHLET, lambda x, nrm: _cos3(x))
                                  if p in ("xmin", "ymax") else
                                  lp.ScalarPatchBC(p, lp.AN_BC_NEUMANN, 0.0)
                                  for p in m.patchNames])
            opts = lp.EllipticOptions(nNonOrthCorr=10, solver=dict(absTol=1e-12, relTol=0.0))
            phi, _ = lp.solve_poisson(m, f, bcs, opts)
            assert abs(_l2(m, phi, _cos3) - r1["l2"]) <= 1e-9 * r1["l2"]
            # pure Neumann: the driver removed the volume-weighted source mean
            mean = float(np.sum(f * vol) / vol.sum())
            assert abs(r2["srcmean"] - mean) <= 1e-12 * max(abs(mean), 1.0)
            bcs = lp.ScalarBCSet([lp.ScalarPatchBC(p, lp.AN_BC_NEUMANN, 0.0) for p in m.patchNames])
            opts = lp.EllipticOptions(nNonOrthCorr=10, useReferenceCell=True, referenceCell=0,
                                      referenceValue=float(_cos3(cc[:1])[0]),
                                      solver=dict(absTol=1e-12, relTol=0.0))
            phi, _ = lp.solve_poisson(m, f - mean, bcs, opts)
            # the reference cell's row is the identity with the exact value
            assert abs(phi[0] - float(_cos3(cc[:1])[0])) <= 1e-10
            assert abs(_l2(m, phi, _cos3) - r2["l2"]) <= 1e-9 * r2["l2"]
    assert np.all(_orders(mixed) > 1.8), mixed
    # pinning one cell to the exact value shifts the whole solution by that
    # cell's discretisation error, a constant of the same order as the error
    assert np.all(_orders(neu) > 1.6), neu


@pytest.mark.gpu
def test_cpp_libscalar_transport_mms(gpu, tmp_path):
    from anabasis_amd import libpoisson as lp
    exe = build_driver(tmp_path)
    central, upwind = [], []
    for n in (8, 16, 32):
        m = lp.Mesh.hexbox(n, n, n)
        d = tmp_path / f"hex{n}"
        m.write_polymesh(d)
        rc = run_driver(exe, "scalar", d, "central", "orth")
        ru = run_driver(exe, "scalar", d, "upwind", "orth")
        assert rc["relres"] <= 1e-6 and ru["relres"] <= 1e-6
        central.append(rc["l2"])
        upwind.append(ru["l2"])
        if n == 8:
            xf, Sf, cc = m.array("xf"), m.array("Sf"), m.array("cc")
            U = lambda x: np.stack([1 + x[:, 1], 0.5 - x[:, 0], np.full(len(x), 0.25)], 1)
            flux = np.sum(U(xf) * Sf, axis=1)
            g = np.pi * np.stack([np.cos(np.pi * cc[:, 0]) * np.sin(np.pi * cc[:, 1]) *
                                  np.sin(np.pi * cc[:, 2]),
                                  np.sin(np.pi * cc[:, 0]) * np.cos(np.pi * cc[:, 1]) *
                                  np.sin(np.pi * cc[:, 2]),
                                  np.sin(np.pi * cc[:, 0]) * np.sin(np.pi * cc[:, 1]) *
                                  np.cos(np.pi * cc[:, 2])], 1)
            Su = np.sum(U(cc) * g, axis=1) + 3 * 0.1 * np.pi ** 2 * _sin3(cc)
            bcs = lp.ScalarBCSet([lp.ScalarPatchBC(p, lp.AN_BC_DIRICHLET, lambda x, nrm: _sin3(x))
                                  for p in m.patchNames])
            phi, _ = lp.solve_steady_scalar_transport(m, flux, np.full(m.nFaces, 0.1), Su,
                                                      np.zeros(m.nCells), bcs, upwind=False,
                                                      nonOrth=False, nNonOrthCorr=0,
                                                      absTol=1e-12)
            assert abs(_l2(m, phi, _sin3) - rc["l2"]) <= 1e-9 * rc["l2"]
    oc, ou = _orders(central), _orders(upwind)
    assert np.all(oc > 1.8), (central, oc)
    # first order, still pre-asymptotic on these meshes (0.66, 0.80 measured)
    assert ou[1] > ou[0] and 0.7 < ou[1] < 1.2, (upwind, ou)
    assert np.all(np.array(central) < np.array(upwind))
