"""generic_simple_v1 driver over the C ABI, plus the runtime boundary
conditions it consumes.

The boundary-condition layer runs in the library (an_bc_*, an_patch_geometry,
following apps/generic_simple_v1/src):
  load_runtime_bc_config                  bc_runtime_config.cu:83-210
  build_patch_geometry_table              patch_geometry.cu:83-140
  apply_bc_specs_to_legacy_face_arrays    velocity_bc_eval.cu:197-290
  coverage table                          main.cu:3178-3233
This module turns spec objects into the .bc grammar and calls it.  The outer
iteration itself runs in libanabasis_hip.so (an_simple_*).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
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

    def line(self) -> str:
        """The spec as a .bc line."""
        t = self.type
        head = f"velocity {self.patchName} {t}"
        if t == "fixed_uniform_vector":
            return head + " " + " ".join(repr(float(v)) for v in self.vector)
        if t == "fixed_normal_speed":
            return f"{head} {float(self.normalSpeed)!r} {self.normalMode}"
        if t == "fixed_flow_rate":
            return f"{head} {float(self.flowRate)!r} {self.normalMode}"
        if t == "parabolic_box_inlet":
            return (f"{head} {float(self.H)!r} {float(self.Umax)!r} {self.coordA} {self.coordB} "
                    f"{self.normalMode}")
        return head


@dataclass
class PressurePatchBC:
    patchName: str
    type: str
    value: float = 0.0

    def line(self) -> str:
        head = f"pressure {self.patchName} {self.type}"
        return f"{head} {float(self.value)!r}" if self.type == "fixed_value" else head


def bc_text(vel: Sequence[VelocityPatchBC], prs: Sequence[PressurePatchBC]) -> str:
    return "".join(s.line() + "\n" for s in list(vel) + list(prs))


class BCConfig:
    """an_bc_t: a parsed runtime BC configuration."""

    def __init__(self, handle):
        self._h = handle

    @staticmethod
    def load(path: str) -> "BCConfig":
        h = C.c_void_p()
        check(lib().an_bc_load(str(path).encode(), C.byref(h)))
        return BCConfig(h)

    @staticmethod
    def parse(text: str, source: str = "<string>") -> "BCConfig":
        h = C.c_void_p()
        check(lib().an_bc_parse(text.encode(), source.encode(), C.byref(h)))
        return BCConfig(h)

    @staticmethod
    def from_specs(vel, prs) -> "BCConfig":
        return BCConfig.parse(bc_text(vel, prs), "<specs>")

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                lib().an_bc_destroy(h)
            except Exception:
                pass
            self._h = None


def _config(vel, prs) -> BCConfig:
    return vel if isinstance(vel, BCConfig) else BCConfig.from_specs(vel, prs)


def patch_geometry_table(mesh) -> List[dict]:
    """patch_geometry.cu:83-140 (an_patch_geometry): area, area-weighted
    centroid, sum S, average normal, planarity, tangents."""
    out = []
    for k, name in enumerate(mesh.patchNames):
        g = np.empty(18)
        check(lib().an_patch_geometry(mesh._h, k, ptr(g)))
        out.append(dict(patchIndex=k, patchName=name, nFaces=int(g[0]), totalArea=g[1],
                        centroid=g[2:5], sumSf=g[5:8], averageNormal=g[8:11], planarity=g[11],
                        tangent1=g[12:15], tangent2=g[15:18]))
    return out


def apply_bc_specs(mesh, vel, prs=None):
    """an_bc_apply (velocity_bc_eval.cu:197-290): per-patch types (1 =
    Dirichlet) and per-face values (nFaces arrays, internal entries 0).
    vel/prs: spec lists, or vel = a BCConfig."""
    cfg = _config(vel, prs)
    nP, nF = len(mesh.patchNames), mesh.nFaces
    tU, tP = np.zeros(nP, np.int32), np.zeros(nP, np.int32)
    uv = [np.zeros(nF) for _ in range(4)]
    check(lib().an_bc_apply(mesh._h, cfg._h, ptr(tU), ptr(tP), *[ptr(a) for a in uv]))
    return tU, tP, uv[0], uv[1], uv[2], uv[3]


def coverage(mesh, vel, prs=None):
    """(table, report, complete) of main.cu:3178-3233 (an_bc_coverage)."""
    cfg = _config(vel, prs)
    out = []
    for which in (0, 1):
        n, comp = C.c_int(), C.c_int()
        check(lib().an_bc_coverage(mesh._h, cfg._h, which, None, 0, C.byref(n), C.byref(comp)))
        buf = C.create_string_buffer(n.value + 1)
        check(lib().an_bc_coverage(mesh._h, cfg._h, which, buf, n.value + 1, C.byref(n),
                                   C.byref(comp)))
        out.append(buf.value.decode())
    return out[0], out[1], bool(comp.value)


def coverage_table(mesh, vel, prs=None) -> List[str]:
    """The coverage lines; raises with the reference's report when a patch
    lacks a velocity or pressure BC (or has two)."""
    table, report, complete = coverage(mesh, vel, prs)
    if not complete:
        raise RuntimeError(report.strip())
    return table.splitlines()


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
        """A cell field, "phi" (one value per face) or "pA" (the last
        pressure matrix's values in CSR order)."""
        if name == "phi":
            n = self.mesh.nFaces
        elif name == "pA":
            own, nei = self.mesh.owner[: self.mesh.nInternalFaces], self.mesh.neigh
            pairs = np.unique(np.minimum(own, nei).astype(np.int64) * self.mesh.nCells
                              + np.maximum(own, nei))
            n = self.mesh.nCells + 2 * pairs.size
        else:
            n = self.mesh.nCells
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
