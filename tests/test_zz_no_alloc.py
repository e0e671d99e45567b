"""an_bicgstab_solve_noalloc: the BiCGSTAB + Jacobi solve of
an_bicgstab_solve (GSV1 solve_momentum_gpu, main.cu:2553-2576) with every
work buffer taken from the library's fixed device scratch array instead of
device allocations.  On the GPU it must give the bits of an_bicgstab_solve
on the same device inputs, on two calls in a row (the second smaller, so
stale scratch contents from the first call would show); without a GPU its
argument checks must reject oversize and NULL inputs before any device
work."""
import ctypes as C

import numpy as np
import pytest

from anabasis_amd import _lib as L
from anabasis_amd import libpoisson as lp


def _call(n, ro, ci, nnz, va, b, x, info):
    return L.lib().an_bicgstab_solve_noalloc(n, ro, ci, nnz, va, b, x, 1000, 0.0, 1e-12,
                                             C.byref(info))


def test_noalloc_rejects_oversize_and_null_before_device_work():
    info = L.SolveInfo()
    dummy = C.c_void_p(1)  # never dereferenced: the checks come first
    n = L.AN_NOALLOC_MAX_ROWS + 1
    st = _call(n, dummy, dummy, 7 * n, dummy, dummy, dummy, info)
    assert st == L.AN_EINVAL
    msg = L.lib().an_last_error().decode()
    assert "exceed the maximum of 400000" in msg and "MiB of scratch" in msg
    st = _call(10, None, dummy, 30, dummy, dummy, dummy, info)
    assert st == L.AN_EINVAL
    st = _call(0, dummy, dummy, 0, dummy, dummy, dummy, info)
    assert st == L.AN_EINVAL


def _system(nx, ny, nz, seed):
    """A non-symmetric momentum-like system: LS central convection +
    diffusion on a jittered hex box."""
    m = lp.Mesh.hexbox(nx, ny, nz, jitter=0.2, seed=seed)
    pat = lp.build_scalar_pattern(m)
    rng = np.random.default_rng(seed)
    U = np.array([1.0, 0.4, -0.3])
    flux = m.array("Sf") @ U
    nb = m.nFaces - m.nInternalFaces
    t = np.ones(nb, np.int32)
    v = np.zeros(nb)
    g = np.zeros((m.nCells, 3))
    vals, _ = lp.assemble_scalar_transport_system(m, flux, np.full(m.nFaces, 0.05),
                                                  np.zeros(m.nCells), np.zeros(m.nCells), t, v,
                                                  g, False, False)
    b = rng.standard_normal(m.nCells)
    return pat, vals, b


def _dev(torch, a, dtype):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype).cuda()


@pytest.mark.gpu
def test_noalloc_matches_bicgstab_bit_for_bit(gpu):
    import torch
    for dims, seed in (((24, 20, 18), 3), ((9, 7, 5), 4)):  # second call smaller
        pat, vals, b = _system(*dims, seed)
        n, nnz = pat.nRows, pat.nnz
        ro = _dev(torch, pat.rowOffsets, torch.int32)
        ci = _dev(torch, pat.cols, torch.int32)
        va = _dev(torch, vals, torch.float64)
        bd = _dev(torch, b, torch.float64)
        xa = torch.zeros(n, dtype=torch.float64, device="cuda")
        xb = torch.zeros(n, dtype=torch.float64, device="cuda")
        ia, ib = L.SolveInfo(), L.SolveInfo()
        torch.cuda.synchronize()
        st = L.lib().an_bicgstab_solve(n, ro.data_ptr(), ci.data_ptr(), nnz, va.data_ptr(),
                                       bd.data_ptr(), xa.data_ptr(), 1000, 0.0, 1e-12,
                                       L.AN_DEVICE, C.byref(ia))
        assert st == L.AN_OK, L.lib().an_last_error()
        st = _call(n, ro.data_ptr(), ci.data_ptr(), nnz, va.data_ptr(), bd.data_ptr(),
                   xb.data_ptr(), ib)
        assert st == L.AN_OK, L.lib().an_last_error()
        torch.cuda.synchronize()
        assert ib.converged and ib.iterations == ia.iterations > 3
        assert ib.finalResNorm == ia.finalResNorm
        assert torch.equal(xa, xb), dims
        # and it is the solution
        import scipy.sparse as sp
        A = sp.csr_matrix((vals, pat.cols, pat.rowOffsets), shape=(n, n))
        x = xb.cpu().numpy()
        assert np.linalg.norm(A @ x - b) <= 1e-10 * np.linalg.norm(b)
