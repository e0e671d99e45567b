"""Python mirror of the libpoisson / libscalar call API over the C ABI.

Names and argument meaning follow LP/include/*.h and
LS/include/scalar_transport_library.h so that tests read like the
reference's own drivers; all numerical work runs in libanabasis_hip.so.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Callable, Dict, List, Optional

import numpy as np

from . import _lib as L
from ._lib import AN_BC_DIRICHLET, AN_BC_NEUMANN, AN_BC_ZERO_GRADIENT, check, f64, i32, lib, ptr


class Mesh:
    """Owns an an_mesh_t (LP/include/mesh.h:12-24 Mesh)."""

    def __init__(self, handle):
        self._h = handle
        info = L.MeshInfo()
        check(lib().an_mesh_get_info(self._h, C.byref(info)))
        self.nPoints = info.nPoints
        self.nFaces = info.nFaces
        self.nInternalFaces = info.nInternalFaces
        self.nCells = info.nCells
        self.nBoundaryFaces = info.nBoundaryFaces
        self.maxNonOrthDeg = info.maxNonOrthDeg
        self.hasGeometry = bool(info.hasGeometry)
        self.patchNames: List[str] = []
        self.patchStartFace: List[int] = []
        self.patchNFaces: List[int] = []
        buf = C.create_string_buffer(256)
        s, n = C.c_int(), C.c_int()
        for k in range(info.nPatches):
            check(lib().an_mesh_get_patch(self._h, k, buf, 256, C.byref(s), C.byref(n)))
            self.patchNames.append(buf.value.decode())
            self.patchStartFace.append(s.value)
            self.patchNFaces.append(n.value)
        self._cache: Dict[str, np.ndarray] = {}

    # constructors ------------------------------------------------------
    @staticmethod
    def read_polymesh(directory: str) -> "Mesh":
        h = C.c_void_p()
        check(lib().an_mesh_read_polymesh(str(directory).encode(), C.byref(h)))
        return Mesh(h)

    @staticmethod
    def hexbox(nx, ny, nz, lx=1.0, ly=1.0, lz=1.0, jitter=0.0, seed=0) -> "Mesh":
        h = C.c_void_p()
        check(lib().an_mesh_gen_hexbox(nx, ny, nz, lx, ly, lz, jitter, seed, C.byref(h)))
        return Mesh(h)

    @staticmethod
    def tetcube(nblocks, length=1.0) -> "Mesh":
        h = C.c_void_p()
        check(lib().an_mesh_gen_tetcube(nblocks, length, C.byref(h)))
        return Mesh(h)

    @staticmethod
    def cylinder(target_cells: int) -> "Mesh":
        h = C.c_void_p()
        check(lib().an_mesh_gen_cylinder(int(target_cells), C.byref(h)))
        return Mesh(h)

    @staticmethod
    def from_arrays(owner, neigh, patches, points=None, face_offsets=None, face_verts=None) -> "Mesh":
        """patches: list of (name, startFace, nFaces)."""
        owner = i32(owner)
        neigh = i32(neigh)
        names = (C.c_char_p * max(len(patches), 1))(*[p[0].encode() for p in patches])
        start = i32([p[1] for p in patches])
        nfc = i32([p[2] for p in patches])
        h = C.c_void_p()
        if points is not None:
            points = f64(points).reshape(-1)
            face_offsets = i32(face_offsets)
            face_verts = i32(face_verts)
            npts = points.size // 3
        else:
            npts = 0
        check(lib().an_mesh_from_arrays(npts, ptr(points), owner.size, ptr(face_offsets),
                                        ptr(face_verts), ptr(owner), neigh.size, ptr(neigh),
                                        len(patches), names, ptr(start), ptr(nfc), C.byref(h)))
        return Mesh(h)

    def write_polymesh(self, directory: str):
        check(lib().an_mesh_write_polymesh(self._h, str(directory).encode()))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                lib().an_mesh_destroy(h)
            except Exception:
                pass
            self._h = None

    # arrays -------------------------------------------------------------
    _SHAPES = {"owner": ("i", "nFaces"), "neigh": ("i", "nInternalFaces"),
               "bPatch": ("i", "nFaces"), "vol": ("d", "nCells"), "Af": ("d", "nFaces"),
               "cc": ("d3", "nCells"), "xf": ("d3", "nFaces"), "nf": ("d3", "nFaces"),
               "Sf": ("d3", "nFaces"), "points": ("d3", "nPoints"),
               "faceOffsets": ("i+1", "nFaces")}

    def array(self, name: str) -> np.ndarray:
        if name in self._cache:
            return self._cache[name]
        if name == "faceVerts":
            off = self.array("faceOffsets")
            out = np.empty(int(off[-1]), dtype=np.int32)
        else:
            kind, n = self._SHAPES[name]
            n = getattr(self, n)
            if kind == "i":
                out = np.empty(n, dtype=np.int32)
            elif kind == "i+1":
                out = np.empty(n + 1, dtype=np.int32)
            elif kind == "d":
                out = np.empty(n, dtype=np.float64)
            else:
                out = np.empty((n, 3), dtype=np.float64)
        check(lib().an_mesh_get_array(self._h, name.encode(), ptr(out)))
        self._cache[name] = out
        return out

    def __getattr__(self, name):
        if name in ("owner", "neigh", "bPatch", "vol", "Af", "cc", "xf", "nf", "Sf"):
            return self.array(name)
        raise AttributeError(name)

    def patch_index(self, name: str) -> int:
        return self.patchNames.index(name)


def read_openfoam_polymesh(directory: str) -> Mesh:
    """LP/src/mesh.cu:160 read_openfoam_polymesh."""
    return Mesh.read_polymesh(directory)


@dataclass
class CSRPattern:
    """LP/include/scalar_elliptic.h:5-14."""
    nRows: int
    nnz: int
    rowOffsets: np.ndarray
    cols: np.ndarray
    diagPos: np.ndarray
    facePP: np.ndarray
    facePN: np.ndarray
    faceNP: np.ndarray
    faceNN: np.ndarray

    @property
    def ncols(self):
        return np.diff(self.rowOffsets)


def build_scalar_pattern(mesh: Mesh) -> CSRPattern:
    """LP/src/scalar_elliptic.cu:16-58, built on the GPU."""
    nC, nIF = mesh.nCells, mesh.nInternalFaces
    ro = np.empty(nC + 1, np.int32)
    dp = np.empty(nC, np.int32)
    check(lib().an_mesh_csr(mesh._h, ptr(ro), None, None, None, None, None, None, L.AN_HOST))
    nnz = int(ro[-1])
    cols = np.empty(nnz, np.int32)
    f = [np.empty(nIF, np.int32) for _ in range(4)]
    check(lib().an_mesh_csr(mesh._h, ptr(ro), ptr(cols), ptr(dp), *[ptr(a) for a in f], L.AN_HOST))
    return CSRPattern(nC, nnz, ro, cols, dp, *f)


def csr_from_faces(nCells, owner, neigh) -> CSRPattern:
    """Same pattern from owner/neighbour only (fixture meshes without points)."""
    owner, neigh = i32(owner), i32(neigh)
    nnz = C.c_int64()
    check(lib().an_csr_from_faces(nCells, owner.size, neigh.size, ptr(owner), ptr(neigh),
                                  C.byref(nnz), None, None, None, None, None, None, None,
                                  L.AN_HOST))
    ro = np.empty(nCells + 1, np.int32)
    cols = np.empty(nnz.value, np.int32)
    dp = np.empty(nCells, np.int32)
    f = [np.empty(neigh.size, np.int32) for _ in range(4)]
    check(lib().an_csr_from_faces(nCells, owner.size, neigh.size, ptr(owner), ptr(neigh),
                                  C.byref(nnz), ptr(ro), ptr(cols), ptr(dp), *[ptr(a) for a in f],
                                  L.AN_HOST))
    return CSRPattern(nCells, nnz.value, ro, cols, dp, *f)


# ----------------------------------------------------------- boundary data
ScalarBCFunc = Callable[[np.ndarray, np.ndarray], np.ndarray]


@dataclass
class ScalarPatchBC:
    """LP/include/bc.h:13-17; evaluator(x[n,3], n[n,3]) -> values[n] or a
    constant."""
    patchName: str
    type: int = AN_BC_DIRICHLET
    evaluator: object = 0.0


@dataclass
class ScalarBCSet:
    patches: List[ScalarPatchBC] = field(default_factory=list)


def build_boundary_face_data(mesh: Mesh, bcset: ScalarBCSet):
    """LP/src/bc.cu:3-35: per boundary face (type, value); raises like the
    reference when a patch has no BC."""
    nIF, nF = mesh.nInternalFaces, mesh.nFaces
    types = np.full(nF - nIF, AN_BC_DIRICHLET, np.int32)
    vals = np.zeros(nF - nIF, np.float64)
    by = {b.patchName: b for b in bcset.patches}
    bP = mesh.bPatch[nIF:]
    if np.any(bP <= 0):
        f = nIF + int(np.argmax(bP <= 0))
        raise RuntimeError(f"Boundary face missing patch index at face {f}")
    xf, nf = mesh.xf[nIF:], mesh.nf[nIF:]
    for k, name in enumerate(mesh.patchNames):
        sel = np.nonzero(bP == k + 1)[0]
        if sel.size == 0:
            continue
        if name not in by:
            raise RuntimeError(f"No BC supplied for patch '{name}'")
        bc = by[name]
        types[sel] = bc.type
        ev = bc.evaluator
        vals[sel] = ev(xf[sel], nf[sel]) if callable(ev) else float(ev)
    return types, vals


# ------------------------------------------------------------ operators
def compute_lsq_gradient(mesh: Mesh, phi, bc_types, bc_vals, mode=L.AN_GRAD_LP) -> np.ndarray:
    """LP/src/gradient.cu:21-83 (mode AN_GRAD_LP) or GSV1
    kernel_apply_lsq_gradient semantics (AN_GRAD_GSV1).  Returns [nC,3]."""
    phi = f64(phi)
    g = np.empty((mesh.nCells, 3), np.float64)
    check(lib().an_lsq_gradient(mesh._h, mode, ptr(i32(bc_types)), ptr(f64(bc_vals)), ptr(phi),
                                ptr(g), L.AN_HOST))
    return g


def assemble_scalar_elliptic_system(mesh, gamma_face, cell_source, bc_types, bc_vals, grad,
                                    include_non_orth=True, use_reference_cell=False, ref_cell=0,
                                    ref_value=0.0, nnz=None):
    """LP/src/scalar_elliptic.cu:60-136 -> (values[nnz], rhs[nC])."""
    if nnz is None:
        nnz = build_scalar_pattern(mesh).nnz
    vals = np.empty(nnz, np.float64)
    rhs = np.empty(mesh.nCells, np.float64)
    check(lib().an_elliptic_assemble(mesh._h, ptr(f64(gamma_face)), ptr(f64(cell_source)),
                                     ptr(i32(bc_types)), ptr(f64(bc_vals)),
                                     ptr(f64(np.asarray(grad).reshape(-1))), int(include_non_orth),
                                     int(use_reference_cell), int(ref_cell), float(ref_value),
                                     ptr(vals), ptr(rhs), L.AN_HOST))
    return vals, rhs


def assemble_scalar_transport_system(mesh, face_flux, gamma_face, Su, Sp, bc_types, bc_vals,
                                     grad, include_non_orth=True, upwind=False, nnz=None):
    """LS/src/scalar_transport_library.cu:250-357 -> (values, rhs)."""
    if nnz is None:
        nnz = build_scalar_pattern(mesh).nnz
    vals = np.empty(nnz, np.float64)
    rhs = np.empty(mesh.nCells, np.float64)
    check(lib().an_transport_assemble(mesh._h, ptr(f64(face_flux)), ptr(f64(gamma_face)),
                                      ptr(f64(Su)), ptr(f64(Sp)), ptr(i32(bc_types)),
                                      ptr(f64(bc_vals)), ptr(f64(np.asarray(grad).reshape(-1))),
                                      int(include_non_orth), int(upwind), ptr(vals), ptr(rhs),
                                      L.AN_HOST))
    return vals, rhs


# --------------------------------------------------------------- solvers
def pcg_options(**kw) -> L.PcgOptions:
    o = L.PcgOptions()
    lib().an_pcg_default_options(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


class PCGSolver:
    """PCG + aggregation AMG on a fixed CSR pattern (the reusable-system API,
    LP/include/hypre_backend.h:38-82)."""

    def __init__(self, rowOffsets, cols, **opts):
        self.ro = i32(rowOffsets)
        self.ci = i32(cols)
        self.n = self.ro.size - 1
        self.opt = pcg_options(**opts)
        h = C.c_void_p()
        check(lib().an_pcg_create(self.n, ptr(self.ro), ptr(self.ci), self.ci.size,
                                  C.byref(self.opt), L.AN_HOST, C.byref(h)))
        self._h = h

    def set_values(self, values, rebuild=True):
        check(lib().an_pcg_set_values(self._h, ptr(f64(values)), int(rebuild), L.AN_HOST))

    def solve(self, b, x0=None, allow_noconv=False):
        x = np.zeros(self.n) if x0 is None else f64(x0).copy()
        info = L.SolveInfo()
        st = lib().an_pcg_solve(self._h, ptr(f64(b)), ptr(x), L.AN_HOST, C.byref(info))
        check(st, allow=(L.AN_ENOCONV,) if allow_noconv else ())
        return x, info

    def residual_history(self):
        """||r0|| and the residual norm after each iteration of the last solve
        (mixed precision: after each refinement step)."""
        cnt = C.c_int()
        check(lib().an_pcg_residual_history(self._h, None, 0, C.byref(cnt)))
        out = np.empty(cnt.value)
        check(lib().an_pcg_residual_history(self._h, ptr(out), cnt.value, C.byref(cnt)))
        return out

    def spmv(self, x):
        y = np.empty(self.n)
        check(lib().an_spmv(self._h, ptr(f64(x)), ptr(y), L.AN_HOST))
        return y

    def amg_apply(self, r):
        z = np.empty(self.n)
        check(lib().an_amg_apply(self._h, ptr(f64(r)), ptr(z), L.AN_HOST))
        return z

    def aggregates(self):
        agg = np.empty(self.n, np.int32)
        nc = C.c_int()
        check(lib().an_amg_aggregates(self._h, ptr(agg), C.byref(nc)))
        return agg, nc.value

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                lib().an_solver_destroy(h)
            except Exception:
                pass
            self._h = None


def bicgstab(rowOffsets, cols, values, b, x0=None, maxIter=1000, relTol=0.0, absTol=1e-10,
             allow_noconv=False):
    ro, ci = i32(rowOffsets), i32(cols)
    n = ro.size - 1
    x = np.zeros(n) if x0 is None else f64(x0).copy()
    info = L.SolveInfo()
    st = lib().an_bicgstab_solve(n, ptr(ro), ptr(ci), ci.size, ptr(f64(values)), ptr(f64(b)),
                                 ptr(x), maxIter, relTol, absTol, L.AN_HOST, C.byref(info))
    check(st, allow=(L.AN_ENOCONV,) if allow_noconv else ())
    return x, info


def bicgstab3(rowOffsets, cols, values, bs, x0s=None, maxIter=1000, relTol=0.0, absTol=1e-10,
              allow_noconv=False):
    """Three right-hand sides sharing one matrix (an_bicgstab_solve3): returns
    ([x0, x1, x2], [info0, info1, info2])."""
    ro, ci = i32(rowOffsets), i32(cols)
    n = ro.size - 1
    bs = [f64(b) for b in bs]
    xs = [np.zeros(n) if x0s is None else f64(x0s[c]).copy() for c in range(3)]
    bp = (C.c_void_p * 3)(*[b.ctypes.data for b in bs])
    xp = (C.c_void_p * 3)(*[x.ctypes.data for x in xs])
    info = (L.SolveInfo * 3)()
    st = lib().an_bicgstab_solve3(n, ptr(ro), ptr(ci), ci.size, ptr(f64(values)), bp, xp,
                                  maxIter, relTol, absTol, L.AN_HOST, info)
    check(st, allow=(L.AN_ENOCONV,) if allow_noconv else ())
    return xs, list(info)


class BatchPCG:
    """Many independent SPD systems solved together (an_batch_*): one
    block-diagonal CSR, one AMG hierarchy, per-problem PCG scalars; the
    iterations run as one persistent launch or a replayed HIP graph, as the
    autotuner chose (set_launch overrides)."""

    def __init__(self, patterns, **opts):
        self.ro = [i32(r) for r, _ in patterns]
        self.ci = [i32(c) for _, c in patterns]
        self.nP = len(patterns)
        self.n = [r.size - 1 for r in self.ro]
        self.opt = pcg_options(**opts)
        rows = np.array(self.n, np.int32)
        rop = (C.c_void_p * self.nP)(*[r.ctypes.data for r in self.ro])
        cip = (C.c_void_p * self.nP)(*[c.ctypes.data for c in self.ci])
        h = C.c_void_p()
        check(lib().an_batch_create(self.nP, ptr(rows), rop, cip, C.byref(self.opt), L.AN_HOST,
                                    C.byref(h)))
        self._h = h

    def set_values(self, values):
        vs = [f64(v) for v in values]
        vp = (C.c_void_p * self.nP)(*[v.ctypes.data for v in vs])
        check(lib().an_batch_set_values(self._h, vp, L.AN_HOST))

    def solve(self, rhs, x0=None):
        bs = [f64(b) for b in rhs]
        xs = [np.zeros(n) if x0 is None else f64(x0[p]).copy() for p, n in enumerate(self.n)]
        bp = (C.c_void_p * self.nP)(*[b.ctypes.data for b in bs])
        xp = (C.c_void_p * self.nP)(*[x.ctypes.data for x in xs])
        info = (L.SolveInfo * self.nP)()
        st = lib().an_batch_solve(self._h, bp, xp, L.AN_HOST, info)
        check(st, allow=(L.AN_ENOCONV,))
        return xs, list(info)

    def use_graph(self, enable=True):
        check(lib().an_batch_use_graph(self._h, int(enable)))

    def set_launch(self, mode):
        """L.AN_BATCH_AUTO / _EAGER / _GRAPH / _PERSISTENT."""
        check(lib().an_batch_set_launch(self._h, int(mode)))

    def tuning_info(self):
        t = L.BatchTuning()
        check(lib().an_batch_get_tuning(self._h, C.byref(t)))
        return dict(rowsPerWave=t.rowsPerWave, rowSeconds=list(t.rowSeconds), launch=t.launch,
                    persistentBlocksPerCU=t.persistentBlocksPerCU,
                    launchSeconds=list(t.launchSeconds), launchesLastSolve=t.launchesLastSolve)

    def tuning(self):
        rows = C.c_int()
        secs = np.zeros(3)
        check(lib().an_batch_tuning(self._h, C.byref(rows), ptr(secs)))
        return rows.value, secs

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                lib().an_batch_destroy(h)
            except Exception:
                pass
            self._h = None


@dataclass
class EllipticOptions:
    """LP/include/poisson_library.h:5-13 (hypre -> PCG/AMG options)."""
    nNonOrthCorr: int = 2
    laplacianScheme: str = "nonorth"
    gradScheme: str = "lsq"
    useReferenceCell: bool = False
    referenceCell: int = 0
    referenceValue: float = 0.0
    solver: dict = field(default_factory=dict)


def solve_scalar_elliptic(mesh: Mesh, gamma_face, cell_source, bcset: ScalarBCSet,
                          opts: EllipticOptions):
    """LP/src/poisson_library.cu:3-47 -> (phi, info)."""
    if opts.gradScheme != "lsq":
        raise RuntimeError("Only gradScheme=lsq is implemented in this build.")
    if opts.laplacianScheme not in ("orth", "nonorth"):
        raise RuntimeError("Use laplacianScheme=orth or laplacianScheme=nonorth.")
    types, vals = build_boundary_face_data(mesh, bcset)
    o = L.EllipticOptions()
    lib().an_elliptic_default_options(C.byref(o))
    o.nNonOrthCorr = opts.nNonOrthCorr
    o.nonOrth = 1 if opts.laplacianScheme == "nonorth" else 0
    o.useReferenceCell = int(opts.useReferenceCell)
    o.referenceCell = opts.referenceCell
    o.referenceValue = opts.referenceValue
    for k, v in opts.solver.items():
        setattr(o.solver, k, v)
    phi = np.empty(mesh.nCells)
    info = L.SolveInfo()
    st = lib().an_solve_scalar_elliptic(mesh._h, ptr(f64(gamma_face)), ptr(f64(cell_source)),
                                        ptr(types), ptr(vals), C.byref(o), ptr(phi), L.AN_HOST,
                                        C.byref(info))
    check(st, allow=(L.AN_ENOCONV,))
    return phi, info


def solve_poisson(mesh: Mesh, cell_source, bcset: ScalarBCSet, opts: EllipticOptions):
    """LP/include/poisson_library.h:27-34 (gamma = 1)."""
    return solve_scalar_elliptic(mesh, np.ones(mesh.nFaces), cell_source, bcset, opts)


def solve_steady_scalar_transport(mesh: Mesh, face_flux, gamma_face, Su, Sp, bcset: ScalarBCSet,
                                  upwind=False, nonOrth=True, nNonOrthCorr=2, maxIter=4000,
                                  relTol=0.0, absTol=1e-10, x0=None):
    """libscalar::solve_steady_scalar_transport
    (LS/src/scalar_transport_library.cu:475-512) -> (phi, info)."""
    types, vals = build_boundary_face_data(mesh, bcset)
    o = L.TransportOptions(int(upwind), int(nonOrth), nNonOrthCorr, maxIter, relTol, absTol)
    phi = np.empty(mesh.nCells)
    info = L.SolveInfo()
    x0p = ptr(f64(x0)) if x0 is not None else None
    st = lib().an_solve_steady_scalar_transport(mesh._h, ptr(f64(face_flux)), ptr(f64(gamma_face)),
                                                ptr(f64(Su)), ptr(f64(Sp)), ptr(types), ptr(vals),
                                                C.byref(o), x0p, ptr(phi), L.AN_HOST,
                                                C.byref(info))
    check(st, allow=(L.AN_ENOCONV,))
    return phi, info
