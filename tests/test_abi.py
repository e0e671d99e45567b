"""C-ABI checks that need no GPU: the library loads, exports every entry point
declared in include/*.h, and its host-only paths (mesh I/O, generators,
argument validation) behave like the reference."""
import ctypes as C
import re

import numpy as np
import pytest

from anabasis_amd import _lib as L
from anabasis_amd.libpoisson import Mesh

from conftest import ROOT


def declared_functions():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        txt = h.read_text()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"\b(an_[a-z0-9_]+)\s*\(", txt):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    lib = L.lib()
    names = declared_functions()
    assert len(names) > 30
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
    assert set(L.SIGNATURES) == names, set(L.SIGNATURES) ^ names


def test_version_and_error_string():
    assert L.lib().an_version() >= 100
    h = C.c_void_p()
    st = L.lib().an_mesh_read_polymesh(b"/nonexistent/polyMesh", C.byref(h))
    assert st == L.AN_EIO
    assert b"Could not open" in L.lib().an_last_error()


def test_hexbox_generator_host_only():
    m = Mesh.hexbox(3, 2, 2)
    assert (m.nCells, m.nInternalFaces) == (12, 2 * 2 * 2 + 3 * 1 * 2 + 3 * 2 * 1)
    assert m.patchNames == ["xmin", "xmax", "ymin", "ymax", "zmin", "zmax"]
    assert sum(m.patchNFaces) == m.nFaces - m.nInternalFaces
    # faces ordered by (owner, neighbour), neighbour > owner
    o, n = m.owner[: m.nInternalFaces], m.neigh
    assert np.all(n > o)
    key = o.astype(np.int64) * m.nCells + n
    assert np.all(np.diff(key) > 0)
    assert np.isclose(m.vol.sum(), 1.0, rtol=1e-14)


def test_polymesh_roundtrip(tmp_path):
    m = Mesh.hexbox(3, 3, 2, jitter=0.2, seed=5)
    d = tmp_path / "polyMesh"
    m.write_polymesh(d)
    r = Mesh.read_polymesh(d)
    assert r.patchNames == m.patchNames
    assert np.array_equal(r.owner, m.owner) and np.array_equal(r.neigh, m.neigh)
    # %.17g round-trips doubles exactly, so the geometry is bit-identical
    for a in ("cc", "vol", "xf", "nf", "Sf", "Af"):
        assert np.array_equal(r.array(a), m.array(a)), a


def test_bad_jitter_rejected():
    h = C.c_void_p()
    assert L.lib().an_mesh_gen_hexbox(2, 2, 2, 1, 1, 1, 0.7, 0, C.byref(h)) == L.AN_EINVAL


def test_topology_only_mesh(cylinder_fixture):
    fx = cylinder_fixture
    m = Mesh.from_arrays(fx["owner"], fx["neigh"], fx["patches"])
    assert m.nCells == 44100 and not m.hasGeometry
    assert m.patchNames == ["patch_1_0", "patch_0_0", "patch_2_0", "patch_5_0", "patch_3_0"]
    with pytest.raises(L.AnError):
        m.array("cc")
