"""Small independent numpy mesh builders used to pin generators/geometry."""
import numpy as np


def hexbox_numpy(nx, ny, nz, lx=1.0, ly=1.0, lz=1.0):
    """Uniform hex box: (points, cell centres, volumes) by closed form."""
    h = np.array([lx / nx, ly / ny, lz / nz])
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    cc = np.stack([(i.T.ravel() + 0.5) * h[0], (j.T.ravel() + 0.5) * h[1],
                   (k.T.ravel() + 0.5) * h[2]], axis=1)
    vol = np.full(nx * ny * nz, h.prod())
    return cc, vol


def bc_arrays(mesh, dirichlet=(), value=None, neumann_value=None):
    """Per-boundary-face type/value arrays: Dirichlet on the named patches."""
    nIF, nF = mesh.nInternalFaces, mesh.nFaces
    t = np.zeros(nF - nIF, np.int32)
    v = np.zeros(nF - nIF)
    bP = mesh.bPatch[nIF:]
    xf, nf = mesh.xf[nIF:], mesh.nf[nIF:]
    for k, name in enumerate(mesh.patchNames):
        sel = bP == k + 1
        if name in dirichlet:
            t[sel] = 1
            if value is not None:
                v[sel] = value(xf[sel])
        elif neumann_value is not None:
            v[sel] = neumann_value(xf[sel], nf[sel])
    return t, v
