"""ctypes wrapper of liboracle.so -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only as the checker (or the timed CPU baseline).  The
product (anabasis_amd) never imports it.
"""
from __future__ import annotations

import ctypes as C
import pathlib

import numpy as np

LIB_PATH = pathlib.Path(__file__).resolve().parent / "liboracle.so"
BC_NEUMANN, BC_DIRICHLET, BC_ZEROGRAD = 0, 1, 2

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} missing: run make -C oracle")
        _lib = C.CDLL(str(LIB_PATH))
        _lib.or_pattern_nnz.restype = C.c_int64
        _lib.or_pcg_amg.restype = C.c_int
        _lib.or_bicgstab_jacobi.restype = C.c_int
    return _lib


def num_threads() -> int:
    """Threads of the oracle's parallel loops (OMP_NUM_THREADS or all cores)."""
    return int(lib().or_num_threads())


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i(a):
    return np.ascontiguousarray(a, dtype=np.int32)


class AmgOptions(C.Structure):
    _fields_ = [("maxIter", C.c_int), ("relTol", C.c_double), ("absTol", C.c_double),
                ("coarseSize", C.c_int), ("maxLevels", C.c_int), ("strength", C.c_double),
                ("pairPasses", C.c_int), ("jacobiOmega", C.c_double), ("sweeps", C.c_int),
                ("kcycle", C.c_int), ("smoother", C.c_int), ("chebLevels", C.c_int),
                ("chebFraction", C.c_double)]


def amg_options(maxIter=4000, relTol=0.0, absTol=1e-12, coarseSize=256, maxLevels=25,
                strength=0.25, pairPasses=3, jacobiOmega=0.85, sweeps=1, kcycle=2, smoother=0,
                chebLevels=1, chebFraction=0.3):
    return AmgOptions(maxIter, relTol, absTol, coarseSize, maxLevels, strength, pairPasses,
                      jacobiOmega, sweeps, kcycle, smoother, chebLevels, chebFraction)


class OMesh:
    """Oracle view of a mesh: plain numpy arrays (interleaved 3-vectors)."""

    def __init__(self, points, face_offsets, face_verts, owner, neigh, patches):
        self.points = _d(points).reshape(-1, 3)
        self.faceOff = _i(face_offsets)
        self.faceVerts = _i(face_verts)
        self.owner = _i(owner)
        self.neigh = _i(neigh)
        self.patches = list(patches)  # (name, start, n)
        self.nFaces = self.owner.size
        self.nInternalFaces = self.neigh.size
        self.nCells = int(max(self.owner.max(initial=-1), self.neigh.max(initial=-1))) + 1
        self.bPatch = np.zeros(self.nFaces, np.int32)
        for k, (_, s, n) in enumerate(self.patches):
            self.bPatch[s:s + n] = k + 1
        self.geometry()
        self.pattern()

    @staticmethod
    def from_product(mesh):
        """Copies the host arrays of an anabasis_amd Mesh."""
        pts = mesh.array("points").copy()
        off = mesh.array("faceOffsets").copy()
        fv = mesh.array("faceVerts").copy()
        patches = list(zip(mesh.patchNames, mesh.patchStartFace, mesh.patchNFaces))
        return OMesh(pts, off, fv, mesh.owner.copy(), mesh.neigh.copy(), patches)

    def geometry(self):
        nC, nF = self.nCells, self.nFaces
        self.cc = np.empty((nC, 3))
        self.vol = np.empty(nC)
        self.xf = np.empty((nF, 3))
        self.nf = np.empty((nF, 3))
        self.Sf = np.empty((nF, 3))
        self.Af = np.empty(nF)
        mx = C.c_double()
        rc = lib().or_mesh_geometry(self.points.shape[0], _p(self.points), nF, _p(self.faceOff),
                                    _p(self.faceVerts), _p(self.owner), self.nInternalFaces,
                                    _p(self.neigh), nC, _p(self.cc), _p(self.vol), _p(self.xf),
                                    _p(self.nf), _p(self.Sf), _p(self.Af), C.byref(mx))
        if rc != 0:
            raise RuntimeError(f"oracle geometry failed ({rc})")
        self.maxNonOrthDeg = mx.value

    def pattern(self):
        self.rowOffsets, self.cols, self.diagPos, self.PP, self.PN, self.NP, self.NN = pattern(
            self.nCells, self.owner, self.neigh)
        self.nnz = self.cols.size


def pattern(nCells, owner, neigh):
    owner, neigh = _i(owner), _i(neigh)
    nF, nIF = owner.size, neigh.size
    nnz = lib().or_pattern_nnz(nCells, nF, nIF, _p(owner), _p(neigh))
    ro = np.empty(nCells + 1, np.int32)
    ci = np.empty(nnz, np.int32)
    dp = np.empty(nCells, np.int32)
    f = [np.empty(nIF, np.int32) for _ in range(4)]
    rc = lib().or_pattern_fill(nCells, nF, nIF, _p(owner), _p(neigh), _p(ro), _p(ci), _p(dp),
                               *[_p(a) for a in f])
    if rc != 0:
        raise RuntimeError("oracle pattern failed")
    return (ro, ci, dp, *f)


def lsq_gradient(m: OMesh, phi, bc_types, bc_vals, mode="lp"):
    g = np.empty((m.nCells, 3))
    phi, bt, bv = _d(phi), _i(bc_types), _d(bc_vals)
    if mode == "gsv1":
        lib().or_lsq_gradient_gsv1(m.nCells, m.nFaces, m.nInternalFaces, _p(m.owner),
                                   _p(m.rowOffsets), _p(m.cols), _p(m.cc), _p(m.xf), _p(bt),
                                   _p(bv), _p(phi), _p(g))
    else:
        lib().or_lsq_gradient_lp(m.nCells, m.nFaces, m.nInternalFaces, _p(m.owner),
                                 _p(m.rowOffsets), _p(m.cols), _p(m.cc), _p(m.xf), _p(m.nf),
                                 _p(bt), _p(bv), _p(phi), _p(g))
    return g


def elliptic_assemble(m: OMesh, gamma, src, bc_types, bc_vals, grad, nonorth=True, use_ref=False,
                      ref_cell=0, ref_value=0.0):
    vals = np.empty(m.nnz)
    rhs = np.empty(m.nCells)
    rc = lib().or_elliptic_assemble(m.nCells, m.nFaces, m.nInternalFaces, _p(m.owner),
                                    _p(m.neigh), _p(m.rowOffsets), _p(m.diagPos), _p(m.PP),
                                    _p(m.PN), _p(m.NP), _p(m.NN), _p(m.cc), _p(m.vol), _p(m.xf),
                                    _p(m.Sf), _p(m.Af), _p(_d(gamma)), _p(_d(src)),
                                    _p(_i(bc_types)), _p(_d(bc_vals)),
                                    _p(_d(np.asarray(grad).reshape(-1))), int(nonorth),
                                    int(use_ref), int(ref_cell), C.c_double(ref_value), _p(vals),
                                    _p(rhs))
    if rc != 0:
        raise RuntimeError("Invalid reference cell")
    return vals, rhs


def transport_assemble(m: OMesh, flux, gamma, Su, Sp, bc_types, bc_vals, grad, nonorth=True,
                       upwind=False):
    vals = np.empty(m.nnz)
    rhs = np.empty(m.nCells)
    lib().or_transport_assemble(m.nCells, m.nFaces, m.nInternalFaces, _p(m.owner), _p(m.neigh),
                                _p(m.diagPos), _p(m.PP), _p(m.PN), _p(m.NP), _p(m.NN), _p(m.cc),
                                _p(m.vol), _p(m.xf), _p(m.Sf), _p(m.Af), _p(_d(flux)),
                                _p(_d(gamma)), _p(_d(Su)), _p(_d(Sp)), _p(_i(bc_types)),
                                _p(_d(bc_vals)), _p(_d(np.asarray(grad).reshape(-1))),
                                int(nonorth), int(upwind), _p(vals), _p(rhs))
    return vals, rhs


def spmv(ro, ci, va, x):
    y = np.empty(len(ro) - 1)
    lib().or_spmv(len(ro) - 1, _p(_i(ro)), _p(_i(ci)), _p(_d(va)), _p(_d(x)), _p(y))
    return y


def pcg_amg(ro, ci, va, b, x0=None, hist_cap=0, **opts):
    n = len(ro) - 1
    x = np.zeros(n) if x0 is None else _d(x0).copy()
    o = amg_options(**opts)
    rr = C.c_double()
    hist = np.zeros(max(hist_cap, 1))
    it = lib().or_pcg_amg(n, _p(_i(ro)), _p(_i(ci)), _p(_d(va)), _p(_d(b)), _p(x), C.byref(o),
                          C.byref(rr), _p(hist) if hist_cap else None, hist_cap)
    return x, it, rr.value, hist


def amg_aggregate(ro, ci, va, **opts):
    n = len(ro) - 1
    agg = np.empty(n, np.int32)
    o = amg_options(**opts)
    nc = lib().or_amg_aggregate(n, _p(_i(ro)), _p(_i(ci)), _p(_d(va)), C.byref(o), _p(agg))
    return agg, nc


def amg_kstep(ro, ci, va, b, **opts):
    n = len(ro) - 1
    x = np.empty(n)
    o = amg_options(**opts)
    nl = lib().or_amg_kstep(n, _p(_i(ro)), _p(_i(ci)), _p(_d(va)), C.byref(o), _p(_d(b)), _p(x))
    return x, nl


def amg_vcycle(ro, ci, va, r, **opts):
    n = len(ro) - 1
    z = np.empty(n)
    o = amg_options(**opts)
    nl = lib().or_amg_vcycle(n, _p(_i(ro)), _p(_i(ci)), _p(_d(va)), C.byref(o), _p(_d(r)), _p(z))
    return z, nl


def pcg_create(ro, ci, **opts):
    """The reusable PCG + AMG handle (or_pcg_create): hierarchy kept between
    solves; pcg_set_values(..., rebuild=False) refreshes only level 0."""
    L = lib()
    L.or_pcg_create.restype = C.c_void_p
    o = amg_options(**opts)
    ro, ci = _i(ro), _i(ci)
    h = L.or_pcg_create(len(ro) - 1, _p(ro), _p(ci), C.byref(o))
    _handles[h] = [ro, ci, None]  # the oracle keeps the arrays' addresses
    return h


def pcg_set_values(h, va, rebuild=True):
    va = _d(va).copy()
    _handles[h][2] = va
    lib().or_pcg_set_values(C.c_void_p(h), _p(va), int(rebuild))


def pcg_solve(h, b, x0=None, relTol=0.0, absTol=1e-12, maxIter=4000):
    ro = _handles[h][0]
    x = np.zeros(len(ro) - 1) if x0 is None else _d(x0).copy()
    rr = C.c_double()
    it = lib().or_pcg_solve(C.c_void_p(h), _p(_d(b)), _p(x), C.c_double(relTol),
                            C.c_double(absTol), int(maxIter), C.byref(rr))
    return x, it, rr.value


def pcg_destroy(h):
    lib().or_pcg_destroy(C.c_void_p(h))
    _handles.pop(h, None)


_handles = {}


def bicgstab(ro, ci, va, b, x0=None, maxIter=1000, relTol=0.0, absTol=1e-10):
    n = len(ro) - 1
    x = np.zeros(n) if x0 is None else _d(x0).copy()
    rr = C.c_double()
    it = lib().or_bicgstab_jacobi(n, _p(_i(ro)), _p(_i(ci)), _p(_d(va)), _p(_d(b)), _p(x),
                                  maxIter, C.c_double(relTol), C.c_double(absTol), C.byref(rr))
    return x, it, rr.value


class SimpleParams(C.Structure):
    _fields_ = [("rho", C.c_double), ("mu", C.c_double), ("nVelNonOrthCorr", C.c_int),
                ("nNonOrthCorr", C.c_int), ("uRelax", C.c_double), ("pRelax", C.c_double),
                ("corrTol", C.c_double), ("tolMass", C.c_double), ("tolVel", C.c_double),
                ("velMaxit", C.c_int), ("velTol", C.c_double), ("velRelTol", C.c_double),
                ("pMaxit", C.c_int), ("pTol", C.c_double), ("pRelTol", C.c_double),
                ("pAmgRebuildEvery", C.c_int), ("amg", AmgOptions)]


class SimpleStats(C.Structure):
    _fields_ = [("massRes", C.c_double), ("duRel", C.c_double), ("dvRel", C.c_double),
                ("dwRel", C.c_double), ("dpRel", C.c_double), ("its", C.c_int * 3),
                ("pcgLast", C.c_int), ("pcgTotal", C.c_int)]


def simple_params_from(gp):
    """Copies an anabasis_amd SimpleParams (ctypes) into the oracle layout."""
    pr = gp.pressure
    amg = amg_options(maxIter=gp.pMaxit, relTol=gp.pRelTol, absTol=gp.pTol,
                      coarseSize=pr.coarseSize, maxLevels=pr.maxLevels, strength=pr.strength,
                      pairPasses=pr.pairPasses, jacobiOmega=pr.jacobiOmega, sweeps=pr.sweeps,
                      kcycle=pr.kcycle, smoother=pr.smoother, chebLevels=pr.chebLevels,
                      chebFraction=pr.chebFraction)
    return SimpleParams(gp.rho, gp.mu, gp.nVelNonOrthCorr, gp.nNonOrthCorr, gp.uRelax,
                        gp.pRelax, gp.corrTol, gp.tolMass, gp.tolVel, gp.velMaxit, gp.velTol,
                        gp.velRelTol, gp.pMaxit, gp.pTol, gp.pRelTol, gp.pAmgRebuildEvery, amg)


def simple_run(m: OMesh, par: SimpleParams, tU, tP, ub, vb, wb, pb, n_iter, u, v, w, p):
    lib().or_simple_run.restype = C.c_int
    u, v, w, p = (_d(a).copy() for a in (u, v, w, p))
    phi = np.empty(m.nFaces)
    rau = np.empty(m.nCells)
    stats = (SimpleStats * max(n_iter, 1))()
    done = lib().or_simple_run(
        m.nCells, m.nFaces, m.nInternalFaces, _p(m.owner), _p(m.neigh), _p(m.bPatch),
        _p(m.rowOffsets), _p(m.cols), _p(m.diagPos), _p(m.PP), _p(m.PN), _p(m.NP), _p(m.NN),
        _p(m.cc), _p(m.vol), _p(m.xf), _p(m.nf), _p(m.Sf), _p(m.Af), C.byref(par),
        _p(_i(tU)), _p(_i(tP)), _p(_d(ub)), _p(_d(vb)), _p(_d(wb)), _p(_d(pb)), n_iter,
        _p(u), _p(v), _p(w), _p(p), _p(phi), _p(rau), stats)
    return dict(u=u, v=v, w=w, p=p, phi=phi, rAU=rau, stats=list(stats)[:abs(done)], done=done)


def patch_forces(m: OMesh, f0, f1, u, v, w, p, mu, sign):
    out = np.empty(15)
    lib().or_patch_forces(_p(m.owner), _p(m.cc), _p(m.xf), _p(m.nf), _p(m.Af), f0, f1,
                          _p(_d(u)), _p(_d(v)), _p(_d(w)), _p(_d(p)), C.c_double(mu), sign,
                          _p(out))
    return out


# ------------------------------------------------- SIMPLE building blocks
# Types are per PATCH (index = bPatch - 1), values per FACE (nFaces entries),
# as in the reference DeviceBC arrays.
def momentum_assemble(m: OMesh, gradQ, uc, vc, wc, gradPcomp, tQ, vQ, tU, vU, tV, vV, tW, vW,
                      rho, mu, corrPsi=1.0):
    vals = np.empty(m.nnz)
    rhs = np.empty(m.nCells)
    lib().or_momentum_assemble(
        m.nCells, m.nFaces, m.nInternalFaces, _p(m.owner), _p(m.neigh), _p(m.bPatch),
        _p(m.diagPos), _p(m.PP), _p(m.PN), _p(m.NP), _p(m.NN), _p(m.cc), _p(m.vol), _p(m.xf),
        _p(m.nf), _p(m.Sf), _p(m.Af), _p(_d(np.asarray(gradQ).reshape(-1))), _p(_d(uc)),
        _p(_d(vc)), _p(_d(wc)), _p(_d(gradPcomp)), _p(_i(tQ)), _p(_d(vQ)), _p(_i(tU)),
        _p(_d(vU)), _p(_i(tV)), _p(_d(vV)), _p(_i(tW)), _p(_d(vW)), C.c_double(rho),
        C.c_double(mu), C.c_double(corrPsi), C.c_int64(m.nnz), _p(vals), _p(rhs))
    return vals, rhs


def momentum_relax(m: OMesh, vals, rhs, qOld, uRelax):
    vals, rhs = _d(vals).copy(), _d(rhs).copy()
    lib().or_momentum_relax(m.nCells, _p(m.diagPos), _p(vals), _p(rhs), _p(_d(qOld)),
                            C.c_double(uRelax))
    return vals, rhs


def extract_rau(m: OMesh, vals):
    out = np.empty(m.nCells)
    lib().or_extract_rau(m.nCells, _p(m.diagPos), _p(_d(vals)), _p(m.vol), _p(out))
    return out


def pressure_assemble(m: OMesh, rAU, tP):
    vals = np.empty(m.nnz)
    lib().or_pressure_assemble(m.nCells, m.nFaces, m.nInternalFaces, _p(m.owner), _p(m.neigh),
                               _p(m.bPatch), _p(m.diagPos), _p(m.PP), _p(m.PN), _p(m.NP),
                               _p(m.NN), _p(m.cc), _p(m.xf), _p(m.nf), _p(m.Af), _p(_d(rAU)),
                               _p(_i(tP)), C.c_int64(m.nnz), _p(vals))
    return vals


def rhiechow_flux(m: OMesh, u, v, w, p, gradP, rAU, tU, ub, tV, vb, tW, wb, rho):
    out = np.empty(m.nFaces)
    lib().or_rhiechow_flux(m.nFaces, m.nInternalFaces, _p(m.owner), _p(m.neigh), _p(m.bPatch),
                           _p(m.cc), _p(m.xf), _p(m.nf), _p(m.Af), _p(_d(u)), _p(_d(v)),
                           _p(_d(w)), _p(_d(p)), _p(_d(np.asarray(gradP).reshape(-1))),
                           _p(_d(rAU)), _p(_i(tU)), _p(_d(ub)), _p(_i(tV)), _p(_d(vb)),
                           _p(_i(tW)), _p(_d(wb)), C.c_double(rho), _p(out))
    return out


def continuity(m: OMesh, phi):
    out = np.empty(m.nCells)
    lib().or_continuity(m.nCells, m.nFaces, m.nInternalFaces, _p(m.owner), _p(m.neigh),
                        _p(_d(phi)), _p(out))
    return out


def pressure_nonorth(m: OMesh, rAU, tP, gradP):
    phi = np.empty(m.nFaces)
    div = np.empty(m.nCells)
    lib().or_pressure_nonorth(m.nCells, m.nFaces, m.nInternalFaces, _p(m.owner), _p(m.neigh),
                              _p(m.bPatch), _p(m.cc), _p(m.xf), _p(m.nf), _p(m.Sf), _p(m.Af),
                              _p(_d(rAU)), _p(_i(tP)), _p(_d(np.asarray(gradP).reshape(-1))),
                              _p(phi), _p(div))
    return phi, div


def correct_flux(m: OMesh, rAU, tP, pb, phiStar, pCorr, phiNo):
    out = np.empty(m.nFaces)
    lib().or_correct_flux(m.nFaces, m.nInternalFaces, _p(m.owner), _p(m.neigh), _p(m.bPatch),
                          _p(m.cc), _p(m.xf), _p(m.nf), _p(m.Af), _p(_d(rAU)), _p(_i(tP)),
                          _p(_d(pb)), _p(_d(phiStar)), _p(_d(pCorr)), _p(_d(phiNo)), _p(out))
    return out


def momentum_rhs_only(m: OMesh, gradQ, uc, vc, wc, gradPcomp, tQ, vQ, tU, vU, tV, vV, tW, vW,
                      rho, mu, corrPsi=1.0):
    rhs = np.empty(m.nCells)
    lib().or_momentum_rhs_only(
        m.nCells, m.nFaces, m.nInternalFaces, _p(m.owner), _p(m.neigh), _p(m.bPatch), _p(m.cc),
        _p(m.vol), _p(m.xf), _p(m.nf), _p(m.Sf), _p(m.Af),
        _p(_d(np.asarray(gradQ).reshape(-1))), _p(_d(uc)), _p(_d(vc)), _p(_d(wc)),
        _p(_d(gradPcomp)), _p(_i(tQ)), _p(_d(vQ)), _p(_i(tU)), _p(_d(vU)), _p(_i(tV)),
        _p(_d(vV)), _p(_i(tW)), _p(_d(vW)), C.c_double(rho), C.c_double(mu), C.c_double(corrPsi),
        _p(rhs))
    return rhs


def momentum_relax_rhs_only(m: OMesh, relaxed_vals, rhs, qOld, uRelax):
    rhs = _d(rhs).copy()
    lib().or_momentum_relax_rhs_only(m.nCells, _p(m.diagPos), _p(_d(relaxed_vals)), _p(rhs),
                                     _p(_d(qOld)), C.c_double(uRelax))
    return rhs
