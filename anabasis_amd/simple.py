"""generic_simple_v1 driver over the C ABI, plus the runtime boundary
conditions it consumes.

Boundary data follow apps/generic_simple_v1/src:
  build_patch_geometry_table   patch_geometry.cu:83-140
  apply_bc_specs_to_legacy_face_arrays   velocity_bc_eval.cu:197-290
  coverage table               main.cu:3178-3233
The outer iteration itself runs in libanabasis_hip.so (an_simple_*).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import List, Sequence, Tuple

import numpy as np

from . import _lib as L
from ._lib import check, f64, i32, lib, ptr

# ------------------------------------------------------------- BC specs ---
VELOCITY_TYPES = ("wall_noslip", "zero_gradient", "fixed_uniform_vector", "fixed_normal_speed",
                  "fixed_flow_rate", "parabolic_box_inlet")
PRESSURE_TYPES = ("zero_gradient", "open", "fixed_value")


@dataclass
class VelocityPatchBC:
    """pipebc::VelocityPatchBCSpec (bc_specs.h)."""
    patchName: str
    type: str
    vector: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    normalSpeed: float = 0.0
    flowRate: float = 0.0
    H: float = 0.0
    Umax: float = 0.0
    coordA: int = 1
    coordB: int = 2
    normalMode: str = "average_patch_normal"  # or local_face_normal


@dataclass
class PressurePatchBC:
    patchName: str
    type: str
    value: float = 0.0


def patch_geometry_table(mesh) -> List[dict]:
    """patch_geometry.cu:83-140: area, area-weighted centroid, sum S,
    average normal sum S/|sum S|, planarity |sum S|/area."""
    out = []
    for k, name in enumerate(mesh.patchNames):
        s, n = mesh.patchStartFace[k], mesh.patchNFaces[k]
        if n < 0 or s < mesh.nInternalFaces or s + n > mesh.nFaces:
            raise RuntimeError(f"Invalid patch face range for patch '{name}'")
        Af = mesh.Af[s:s + n]
        area = 0.0
        cacc = np.zeros(3)
        ssum = np.zeros(3)
        for f in range(s, s + n):  # serial, face order
            area += mesh.Af[f]
            cacc = cacc + mesh.Af[f] * mesh.xf[f]
            ssum = ssum + mesh.Sf[f]
        nrm = float(np.sqrt(ssum @ ssum))
        out.append(dict(patchIndex=k, patchName=name, nFaces=n, totalArea=area,
                        centroid=cacc / area if area > 1e-30 else np.zeros(3), sumSf=ssum,
                        averageNormal=ssum / nrm if nrm > 1e-30 else np.zeros(3),
                        planarity=nrm / area if area > 1e-30 else 0.0))
        del Af
    return out


def _parabolic(spec: VelocityPatchBC, x, nrm):
    H, Um = spec.H, spec.Umax
    if H <= 0.0:
        raise RuntimeError(f"ParabolicBoxInlet has non-positive H on patch '{spec.patchName}'")
    a = np.clip(x[:, spec.coordA], 0.0, H)
    b = np.clip(x[:, spec.coordB], 0.0, H)
    H2 = H * H
    H4 = H2 * H2
    mag = 16.0 * Um * a * b * (H - a) * (H - b) / H4
    return -mag[:, None] * nrm  # inflow along the inward normal (velocity_bc_eval.cu:97-99)


def apply_bc_specs(mesh, vel: Sequence[VelocityPatchBC], prs: Sequence[PressurePatchBC]):
    """velocity_bc_eval.cu:197-290 -> per-patch types (1 = Dirichlet) and
    per-face values (nFaces arrays, internal entries 0)."""
    nP, nF = len(mesh.patchNames), mesh.nFaces
    tU = np.zeros(nP, np.int32)  # Neumann unless set (main.cu:3262-3265)
    tP = np.zeros(nP, np.int32)
    uv = [np.zeros(nF) for _ in range(3)]
    pv = np.zeros(nF)
    geo = patch_geometry_table(mesh)
    for spec in vel:
        if spec.type not in VELOCITY_TYPES:
            raise RuntimeError(f"unknown velocity BC type '{spec.type}'")
        k = mesh.patch_index(spec.patchName)
        s, n = mesh.patchStartFace[k], mesh.patchNFaces[k]
        if spec.type == "zero_gradient":
            tU[k] = 0
            continue
        tU[k] = 1
        xf, nf = mesh.xf[s:s + n], mesh.nf[s:s + n]
        g = geo[k]
        if spec.normalMode == "average_patch_normal":
            an = g["averageNormal"]
            if np.linalg.norm(an) <= 1e-30:
                raise RuntimeError(f"Cannot normalize vector for AveragePatchNormal on patch "
                                   f"'{spec.patchName}'")
            nrm = np.broadcast_to(an / np.linalg.norm(an), (n, 3))
        else:
            nrm = nf / np.linalg.norm(nf, axis=1)[:, None]
        if spec.type == "wall_noslip":
            U = np.zeros((n, 3))
        elif spec.type == "fixed_uniform_vector":
            U = np.broadcast_to(np.array(spec.vector, float), (n, 3))
        elif spec.type == "parabolic_box_inlet":
            U = _parabolic(spec, xf, nrm)
        else:
            if spec.type == "fixed_flow_rate":
                if spec.normalMode == "average_patch_normal" and g["planarity"] < 0.98:
                    raise RuntimeError("FixedFlowRate with AveragePatchNormal requires an "
                                       f"approximately planar patch. Patch '{spec.patchName}'")
                if g["totalArea"] <= 1e-30:
                    raise RuntimeError(f"Patch '{spec.patchName}' has near-zero area")
                un = spec.flowRate / g["totalArea"]
            else:
                un = spec.normalSpeed
            U = un * nrm
        for d in range(3):
            uv[d][s:s + n] = U[:, d]
    for spec in prs:
        if spec.type not in PRESSURE_TYPES:
            raise RuntimeError(f"unknown pressure BC type '{spec.type}'")
        k = mesh.patch_index(spec.patchName)
        s, n = mesh.patchStartFace[k], mesh.patchNFaces[k]
        if spec.type == "fixed_value":
            tP[k] = 1
            pv[s:s + n] = spec.value
        else:
            tP[k] = 0
    return tU, tP, uv[0], uv[1], uv[2], pv


def coverage_table(mesh, vel, prs) -> List[str]:
    """main.cu:3178-3233: one line per patch; raises when a patch lacks a
    velocity or pressure BC, or has two."""
    for specs, what in ((vel, "velocity"), (prs, "pressure")):
        names = [s.patchName for s in specs]
        dup = sorted({n for n in names if names.count(n) > 1})
        if dup:
            raise RuntimeError(f"Duplicate {what} BC specs found: {dup}")
    lines, missV, missP = [], [], []
    for name in mesh.patchNames:
        hv = any(s.patchName == name for s in vel)
        hp = any(s.patchName == name for s in prs)
        lines.append(f"  {name:<32s} velocity={'yes' if hv else 'NO'} "
                     f"pressure={'yes' if hp else 'NO'}")
        if not hv:
            missV.append(name)
        if not hp:
            missP.append(name)
    if missV or missP:
        raise RuntimeError("incomplete runtime BC coverage: missing velocity "
                           f"{missV}, pressure {missP}")
    return lines


def cylinder_case_bcs():
    """Inline BC lines of cases/cylinder.case:55-65."""
    vel = [VelocityPatchBC("patch_0_0", "wall_noslip"),
           VelocityPatchBC("patch_5_0", "wall_noslip"),
           VelocityPatchBC("patch_3_0", "wall_noslip"),
           VelocityPatchBC("patch_1_0", "parabolic_box_inlet", H=0.41, Umax=0.45, coordA=1,
                           coordB=2, normalMode="average_patch_normal"),
           VelocityPatchBC("patch_2_0", "zero_gradient")]
    prs = [PressurePatchBC("patch_0_0", "zero_gradient"),
           PressurePatchBC("patch_5_0", "zero_gradient"),
           PressurePatchBC("patch_3_0", "zero_gradient"),
           PressurePatchBC("patch_1_0", "zero_gradient"),
           PressurePatchBC("patch_2_0", "fixed_value", 0.0)]
    return vel, prs


# --------------------------------------------------------------- driver ---
def default_params(**kw) -> L.SimpleParams:
    p = L.SimpleParams()
    lib().an_simple_default_params(C.byref(p))
    for k, v in kw.items():
        if k in ("pressure",):
            for kk, vv in v.items():
                setattr(p.pressure, kk, vv)
        else:
            setattr(p, k, v)
    return p


class Simple:
    """an_simple_t: the SIMPLE state on the GPU."""

    def __init__(self, mesh, params: L.SimpleParams, vel, prs, inlet_patch="patch_1_0",
                 bc=None):
        """vel/prs: BC specs; or bc = (tU, tP, u, v, w, p) per-patch types and
        per-face values given directly (vel and prs are then ignored)."""
        self.mesh = mesh  # keeps the mesh handle alive
        self.params = params
        self.bc = tuple(bc) if bc is not None else apply_bc_specs(mesh, vel, prs)
        tU, tP, ub, vb, wb, pb = self.bc
        inlet = mesh.patch_index(inlet_patch) if inlet_patch in mesh.patchNames else -1
        h = C.c_void_p()
        check(lib().an_simple_create(mesh._h, C.byref(params), ptr(i32(tU)), ptr(i32(tP)),
                                     ptr(f64(ub)), ptr(f64(vb)), ptr(f64(wb)), ptr(f64(pb)),
                                     inlet, C.byref(h)))
        self._h = h

    def iterate(self, n=1):
        st = L.SimpleStats()
        conv = C.c_int()
        check(lib().an_simple_iterate(self._h, n, C.byref(st), C.byref(conv)))
        return st, bool(conv.value)

    def field(self, name):
        n = self.mesh.nFaces if name == "phi" else self.mesh.nCells
        out = np.empty(n)
        check(lib().an_simple_get_field(self._h, name.encode(), ptr(out), L.AN_HOST))
        return out

    def set_field(self, name, arr):
        check(lib().an_simple_set_field(self._h, name.encode(), ptr(f64(arr)), L.AN_HOST))

    def patch_forces(self, patch_name, normal_sign=-1):
        out = np.empty(15)
        check(lib().an_simple_patch_forces(self._h, self.mesh.patch_index(patch_name),
                                           normal_sign, ptr(out)))
        return out

    def timings(self):
        out = np.empty(8)
        check(lib().an_simple_timings(self._h, ptr(out)))
        return dict(zip(("mom_asm", "mom_solve", "p_asm_setup", "p_solve", "other", "total",
                         "p_pcg_iterations", "p_solves"), out.tolist()))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                lib().an_simple_destroy(h)
            except Exception:
                pass
            self._h = None
