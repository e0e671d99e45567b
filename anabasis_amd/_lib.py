"""ctypes binding of libanabasis_hip.so (the C ABI in include/anabasis_hip.h).

The shared library is built in-tree by ``__graft_entry__.build()`` (or
``make -C anabasis_amd/csrc``).  Loading fails loudly when it is missing: there
is no CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes as C
import os
import pathlib

import numpy as np

LIB_PATH = pathlib.Path(__file__).resolve().parent / "libanabasis_hip.so"

AN_OK, AN_EINVAL, AN_ENOCONV, AN_EHIP, AN_EIO, AN_ENOMEM, AN_EDIVERGED, AN_EINTERNAL = range(8)
AN_HOST, AN_DEVICE = 0, 1
AN_BC_NEUMANN, AN_BC_DIRICHLET, AN_BC_ZERO_GRADIENT = 0, 1, 2
AN_GRAD_GSV1, AN_GRAD_LP = 0, 1
AN_NOALLOC_MAX_ROWS = 400000  # include/anabasis_hip.h

_STATUS = {
    AN_EINVAL: "AN_EINVAL", AN_ENOCONV: "AN_ENOCONV", AN_EHIP: "AN_EHIP", AN_EIO: "AN_EIO",
    AN_ENOMEM: "AN_ENOMEM", AN_EDIVERGED: "AN_EDIVERGED", AN_EINTERNAL: "AN_EINTERNAL",
}


class AnError(RuntimeError):
    """Raised for any non-OK status; ``code`` holds the an_status."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"{_STATUS.get(code, code)}: {msg}")
        self.code = code


class MeshInfo(C.Structure):
    _fields_ = [("nPoints", C.c_int), ("nFaces", C.c_int), ("nInternalFaces", C.c_int),
                ("nCells", C.c_int), ("nPatches", C.c_int), ("nBoundaryFaces", C.c_int),
                ("nnz", C.c_int64), ("maxNonOrthDeg", C.c_double), ("hasGeometry", C.c_int)]


class PcgOptions(C.Structure):
    _fields_ = [("maxIter", C.c_int), ("relTol", C.c_double), ("absTol", C.c_double),
                ("precision", C.c_int), ("coarseSize", C.c_int), ("maxLevels", C.c_int),
                ("strength", C.c_double), ("pairPasses", C.c_int), ("jacobiOmega", C.c_double),
                ("innerRelTol", C.c_double), ("profile", C.c_int), ("sweeps", C.c_int),
                ("kcycle", C.c_int), ("graphs", C.c_int), ("rhsProjection", C.c_int),
                ("smoother", C.c_int), ("chebLevels", C.c_int), ("chebFraction", C.c_double)]


class SolveInfo(C.Structure):
    _fields_ = [("iterations", C.c_int), ("finalRelResNorm", C.c_double),
                ("finalResNorm", C.c_double), ("converged", C.c_int), ("levels", C.c_int),
                ("setupSeconds", C.c_double), ("solveSeconds", C.c_double),
                ("refinementSteps", C.c_int), ("spmvCalls", C.c_int),
                ("spmvAvgSeconds", C.c_double), ("vcycleCalls", C.c_int),
                ("vcycleAvgSeconds", C.c_double), ("spmvIndexBytes", C.c_int)]


class EllipticOptions(C.Structure):
    _fields_ = [("nNonOrthCorr", C.c_int), ("nonOrth", C.c_int), ("useReferenceCell", C.c_int),
                ("referenceCell", C.c_int), ("referenceValue", C.c_double),
                ("solver", PcgOptions)]


class TransportOptions(C.Structure):
    _fields_ = [("upwind", C.c_int), ("nonOrth", C.c_int), ("nNonOrthCorr", C.c_int),
                ("maxIter", C.c_int), ("relTol", C.c_double), ("absTol", C.c_double)]


class SimpleParams(C.Structure):
    _fields_ = [("rho", C.c_double), ("mu", C.c_double), ("nVelNonOrthCorr", C.c_int),
                ("nNonOrthCorr", C.c_int), ("uRelax", C.c_double), ("pRelax", C.c_double),
                ("corrTol", C.c_double), ("tolMass", C.c_double), ("tolVel", C.c_double),
                ("velMaxit", C.c_int), ("velTol", C.c_double), ("velRelTol", C.c_double),
                ("pMaxit", C.c_int), ("pTol", C.c_double), ("pRelTol", C.c_double),
                ("pAmgRebuildEvery", C.c_int), ("precision", C.c_int),
                ("fuseMomentumRhs", C.c_int), ("pressure", PcgOptions)]


class BatchTuning(C.Structure):
    _fields_ = [("rowsPerWave", C.c_int), ("rowSeconds", C.c_double * 5), ("launch", C.c_int),
                ("persistentBlocksPerCU", C.c_int), ("launchSeconds", C.c_double * 4),
                ("launchesLastSolve", C.c_int)]


AN_BATCH_AUTO, AN_BATCH_EAGER, AN_BATCH_GRAPH, AN_BATCH_PERSISTENT = 0, 1, 2, 3


class SimpleStats(C.Structure):
    _fields_ = [("step", C.c_int), ("massRes", C.c_double), ("duRel", C.c_double),
                ("dvRel", C.c_double), ("dwRel", C.c_double), ("dpRel", C.c_double),
                ("bicgstabIts", C.c_int * 3), ("pcgLast", C.c_int), ("pcgTotal", C.c_int),
                ("seconds", C.c_double)]


_P = C.c_void_p
_I = C.c_int
_D = C.c_double
_pI = C.POINTER(C.c_int)
_pD = C.POINTER(C.c_double)

# name -> (restype, argtypes); every symbol declared in include/anabasis_hip.h
SIGNATURES = {
    "an_last_error": (C.c_char_p, []),
    "an_version": (_I, []),
    "an_set_device": (_I, [_I]),
    "an_mesh_read_polymesh": (_I, [C.c_char_p, C.POINTER(_P)]),
    "an_mesh_from_arrays": (_I, [_I, _P, _I, _P, _P, _P, _I, _P, _I, _P, _P, _P, C.POINTER(_P)]),
    "an_mesh_write_polymesh": (_I, [_P, C.c_char_p]),
    "an_mesh_gen_hexbox": (_I, [_I, _I, _I, _D, _D, _D, _D, C.c_uint64, C.POINTER(_P)]),
    "an_mesh_gen_tetcube": (_I, [_I, _D, C.POINTER(_P)]),
    "an_mesh_gen_cylinder": (_I, [C.c_longlong, C.POINTER(_P)]),
    "an_mesh_destroy": (_I, [_P]),
    "an_mesh_get_info": (_I, [_P, C.POINTER(MeshInfo)]),
    "an_mesh_get_array": (_I, [_P, C.c_char_p, _P]),
    "an_mesh_get_patch": (_I, [_P, _I, C.c_char_p, _I, _pI, _pI]),
    "an_mesh_csr": (_I, [_P, _P, _P, _P, _P, _P, _P, _P, _I]),
    "an_csr_from_faces": (_I, [_I, _I, _I, _P, _P, C.POINTER(C.c_int64), _P, _P, _P, _P, _P, _P,
                               _P, _I]),
    "an_lsq_gradient": (_I, [_P, _I, _P, _P, _P, _P, _I]),
    "an_elliptic_assemble": (_I, [_P, _P, _P, _P, _P, _P, _I, _I, _I, _D, _P, _P, _I]),
    "an_transport_assemble": (_I, [_P, _P, _P, _P, _P, _P, _P, _P, _I, _I, _P, _P, _I]),
    "an_pcg_default_options": (None, [C.POINTER(PcgOptions)]),
    "an_pcg_create": (_I, [_I, _P, _P, C.c_int64, C.POINTER(PcgOptions), _I, C.POINTER(_P)]),
    "an_pcg_set_values": (_I, [_P, _P, _I, _I]),
    "an_pcg_solve": (_I, [_P, _P, _P, _I, C.POINTER(SolveInfo)]),
    "an_pcg_set_stream": (_I, [_P, _P]),
    "an_pcg_residual_history": (_I, [_P, _P, _I, C.POINTER(C.c_int)]),
    "an_spmv": (_I, [_P, _P, _P, _I]),
    "an_amg_apply": (_I, [_P, _P, _P, _I]),
    "an_amg_aggregates": (_I, [_P, _P, _pI]),
    "an_solver_destroy": (_I, [_P]),
    "an_bicgstab_solve": (_I, [_I, _P, _P, C.c_int64, _P, _P, _P, _I, _D, _D, _I,
                               C.POINTER(SolveInfo)]),
    "an_bicgstab_solve3": (_I, [_I, _P, _P, C.c_int64, _P, _P, _P, _I, _D, _D, _I,
                                C.POINTER(SolveInfo)]),
    "an_bicgstab_solve_noalloc": (_I, [_I, _P, _P, C.c_int64, _P, _P, _P, _I, _D, _D,
                                       C.POINTER(SolveInfo)]),
    "an_elliptic_default_options": (None, [C.POINTER(EllipticOptions)]),
    "an_solve_scalar_elliptic": (_I, [_P, _P, _P, _P, _P, C.POINTER(EllipticOptions), _P, _I,
                                      C.POINTER(SolveInfo)]),
    "an_transport_default_options": (None, [C.POINTER(TransportOptions)]),
    "an_solve_steady_scalar_transport": (_I, [_P, _P, _P, _P, _P, _P, _P,
                                              C.POINTER(TransportOptions), _P, _P, _I,
                                              C.POINTER(SolveInfo)]),
    "an_batch_create": (_I, [_I, _P, _P, _P, C.POINTER(PcgOptions), _I, C.POINTER(_P)]),
    "an_batch_set_values": (_I, [_P, _P, _I]),
    "an_batch_solve": (_I, [_P, _P, _P, _I, C.POINTER(SolveInfo)]),
    "an_batch_tuning": (_I, [_P, _pI, _P]),
    "an_batch_use_graph": (_I, [_P, _I]),
    "an_batch_set_launch": (_I, [_P, _I]),
    "an_batch_get_tuning": (_I, [_P, C.c_void_p]),
    "an_batch_destroy": (_I, [_P]),
    "an_transport_adjoint": (_I, [_P, _P, _P, _P, _P, _P, _P, C.POINTER(TransportOptions), _P,
                                  _P, _P, _P, _P, _I, C.POINTER(SolveInfo)]),
    "an_face_gamma_from_cells": (_I, [_P, _P, _P, _I]),
    "an_dgamma_to_cells": (_I, [_P, _P, _P, _I]),
    "an_lsq_gradient_adjoint": (_I, [_P, _P, _P, _P, _P, _P, _I]),
    "an_simple_default_params": (None, [C.POINTER(SimpleParams)]),
    "an_simple_create": (_I, [_P, C.POINTER(SimpleParams), _P, _P, _P, _P, _P, _P, _I,
                              C.POINTER(_P)]),
    "an_bc_load": (_I, [C.c_char_p, C.POINTER(_P)]),
    "an_bc_parse": (_I, [C.c_char_p, C.c_char_p, C.POINTER(_P)]),
    "an_bc_destroy": (_I, [_P]),
    "an_patch_geometry": (_I, [_P, _I, _P]),
    "an_bc_coverage": (_I, [_P, _P, _I, C.c_char_p, _I, _pI, _pI]),
    "an_bc_apply": (_I, [_P, _P, _P, _P, _P, _P, _P, _P]),
    "an_simple_iterate": (_I, [_P, _I, C.POINTER(SimpleStats), _pI]),
    "an_simple_get_field": (_I, [_P, C.c_char_p, _P, _I]),
    "an_simple_set_field": (_I, [_P, C.c_char_p, _P, _I]),
    "an_simple_patch_forces": (_I, [_P, _I, _I, _P]),
    "an_simple_timings": (_I, [_P, _P]),
    "an_simple_destroy": (_I, [_P]),
}

_lib = None


def _load_torch_runtime_first():
    """PyTorch's ROCm wheel carries its own HIP and HSA runtimes under
    unversioned file names (torch/lib/libamdhip64.so, libhsa-runtime64.so).
    Loaded after this library (which links /opt/rocm's libhsa-runtime64.so.1),
    torch brings a second HSA runtime into the process, and whichever of the
    two initialises second finds no GPU.  Loaded first, torch's HSA runtime
    carries the soname this library asks for and is shared by both.  So when
    torch is installed it is imported before the library is opened."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def lib():
    """Loads the library once; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} is missing: run __graft_entry__.build() "
                              "(make -C anabasis_amd/csrc)")
        _load_torch_runtime_first()
        L = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name, None)
            if fn is None:  # reported by tests/test_abi.py
                continue
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(status: int, allow=()) -> int:
    if status != AN_OK and status not in allow:
        msg = lib().an_last_error().decode(errors="replace")
        raise AnError(status, msg)
    return status


def ptr(a):
    """Pointer to a contiguous numpy array (or None)."""
    if a is None:
        return None
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("array must be C-contiguous")
    return a.ctypes.data_as(C.c_void_p)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def gpu_available() -> bool:
    if os.environ.get("ANABASIS_FORCE_NO_GPU"):
        return False
    try:
        import torch  # noqa: F401  (device presence only)
        return bool(torch.cuda.is_available())
    except Exception:  # pragma: no cover
        return False
