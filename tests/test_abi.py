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


def _mesh_arrays(m):
    return {k: m.array(k) for k in ("points", "faceOffsets", "faceVerts", "owner", "neigh",
                                    "bPatch", "cc", "vol", "xf", "nf", "Sf", "Af")}


def test_binary_mesh_cache(tmp_path, monkeypatch):
    """read_polymesh writes <dir>/.anabasis_mesh.bin after parsing and reads it
    back while the ASCII files are unchanged: the same arrays bit for bit; a
    rewritten, edited or corrupted mesh directory falls back to the parse."""
    import os
    monkeypatch.delenv("ANABASIS_MESH_CACHE", raising=False)
    d = tmp_path / "polyMesh"
    m0 = Mesh.hexbox(5, 4, 3, 1.0, 0.8, 0.6, jitter=0.2, seed=1)
    m0.write_polymesh(d)
    cache = d / ".anabasis_mesh.bin"
    assert not cache.exists()
    a = _mesh_arrays(Mesh.read_polymesh(d))
    assert cache.exists()
    b = _mesh_arrays(Mesh.read_polymesh(d))
    ref = _mesh_arrays(m0)
    for k in ref:
        assert np.array_equal(a[k], ref[k]) and np.array_equal(b[k], ref[k]), k
    # rewriting the directory drops the cache; the new mesh is read
    m1 = Mesh.hexbox(5, 4, 3, 1.0, 0.8, 0.6, jitter=0.2, seed=2)
    m1.write_polymesh(d)
    assert not cache.exists()
    c = _mesh_arrays(Mesh.read_polymesh(d))
    assert np.array_equal(c["cc"], m1.array("cc")) and not np.array_equal(c["cc"], a["cc"])
    # an edit of the same size (points moved by hand): the stamps differ
    pts = (d / "points").read_text()
    first = pts.index("(", pts.index("(") + 1)
    edited = pts[:first] + pts[first:].replace("0.8", "0.9", 1)
    assert len(edited) == len(pts)
    (d / "points").write_text(edited)
    e = _mesh_arrays(Mesh.read_polymesh(d))
    assert not np.array_equal(e["points"], c["points"])
    # a truncated cache is ignored
    blob = cache.read_bytes()
    cache.write_bytes(blob[: len(blob) // 2])
    f = _mesh_arrays(Mesh.read_polymesh(d))
    assert np.array_equal(f["points"], e["points"]) and np.array_equal(f["cc"], e["cc"])
    # ANABASIS_MESH_CACHE=0: neither read nor written
    os.remove(cache)
    monkeypatch.setenv("ANABASIS_MESH_CACHE", "0")
    Mesh.read_polymesh(d)
    assert not cache.exists()
