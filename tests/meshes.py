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


def refined_interface_arrays(ncx=3, ncy=3, ncz=3, nfx=3, h=0.2, shear=(0.1, 0.05)):
    """A polyhedral mesh with a 2:1 non-conformal interface, as polyMesh
    arrays for Mesh.from_arrays: a coarse block of ncx x ncy x ncz cubes of
    side h, and to its +x side a block of cubes of side h/2 (nfx deep, twice
    as many across).  Each coarse cell on the interface has its +x face
    split into the four faces of its fine neighbours, so it has up to nine
    neighbours (ten nonzeros per matrix row, nine least-squares terms).
    Internal faces are in (owner, neighbour) order with the normal from
    owner to neighbour; boundary faces are grouped by patch xmin, xmax, ymin,
    ymax, zmin, zmax.  The points are sheared (x += s0 y, y += s1 z), which
    keeps every face planar and every cell closed.

    Returns (owner, neigh, patches, points, face_offsets, face_verts)."""
    NX, NY, NZ = 2 * ncx + nfx, 2 * ncy, 2 * ncz  # extent in half-spacing units
    boxes = []  # cells as [x0, x1) x [y0, y1) x [z0, z1) in half-spacing units
    for k in range(ncz):
        for j in range(ncy):
            for i in range(ncx):
                boxes.append((2 * i, 2 * i + 2, 2 * j, 2 * j + 2, 2 * k, 2 * k + 2))
    for k in range(NZ):
        for j in range(NY):
            for i in range(nfx):
                x = 2 * ncx + i
                boxes.append((x, x + 1, j, j + 1, k, k + 1))
    at = np.full((NX, NY, NZ), -1, np.int64)  # cell holding each unit cube
    for c, (x0, x1, y0, y1, z0, z1) in enumerate(boxes):
        at[x0:x1, y0:y1, z0:z1] = c
    pts, pid = [], {}

    def point(p):
        if p not in pid:
            pid[p] = len(pts)
            pts.append(p)
        return pid[p]

    def rect(axis, plane, a0, a1, b0, b1):
        # corners of the rectangle {axis = plane} x [a0, a1] x [b0, b1], in
        # right-handed order around +axis
        u, v = [(1, 2), (2, 0), (0, 1)][axis]
        corners = []
        for a, b in ((a0, b0), (a1, b0), (a1, b1), (a0, b1)):
            q = [0, 0, 0]
            q[axis], q[u], q[v] = plane, a, b
            corners.append(tuple(q))
        return corners

    internal, bnd = [], {n: [] for n in ("xmin", "xmax", "ymin", "ymax", "zmin", "zmax")}
    for c, bx in enumerate(boxes):
        lo, hi = bx[0::2], bx[1::2]
        for axis in range(3):
            u, v = [(1, 2), (2, 0), (0, 1)][axis]
            for side in (0, 1):
                plane = hi[axis] if side else lo[axis]
                across = plane if side else plane - 1
                extent = (NX, NY, NZ)[axis]
                groups = {}
                for a in range(lo[u], hi[u]):
                    for b in range(lo[v], hi[v]):
                        q = [0, 0, 0]
                        q[axis], q[u], q[v] = across, a, b
                        n = int(at[tuple(q)]) if 0 <= across < extent else -1
                        g = groups.setdefault(n, [a, a + 1, b, b + 1])
                        g[0], g[1] = min(g[0], a), max(g[1], a + 1)
                        g[2], g[3] = min(g[2], b), max(g[3], b + 1)
                for n, (a0, a1, b0, b1) in groups.items():
                    corners = rect(axis, plane, a0, a1, b0, b1)
                    if n < 0:  # boundary: outward normal
                        if not side:
                            corners = corners[::-1]
                        name = "xyz"[axis] + ("max" if side else "min")
                        bnd[name].append((c, corners))
                    elif side:  # each internal face once, from its lower side
                        own, nei = min(c, n), max(c, n)
                        if own != c:
                            corners = corners[::-1]  # normal from owner to neighbour
                        internal.append((own, nei, corners))
    internal.sort(key=lambda f: (f[0], f[1]))
    owner, neigh, faces, patches = [], [], [], []
    for own, nei, corners in internal:
        owner.append(own)
        neigh.append(nei)
        faces.append(corners)
    for name in ("xmin", "xmax", "ymin", "ymax", "zmin", "zmax"):
        lst = sorted(bnd[name], key=lambda f: f[0])
        patches.append((name, len(owner), len(lst)))
        for c, corners in lst:
            owner.append(c)
            faces.append(corners)
    offsets, verts = [0], []
    for corners in faces:
        verts += [point(p) for p in corners]
        offsets.append(len(verts))
    P = np.array(pts, float) * (h / 2)
    P[:, 0] += shear[0] * P[:, 1]
    P[:, 1] += shear[1] * P[:, 2]
    return (np.array(owner, np.int32), np.array(neigh, np.int32), patches, P,
            np.array(offsets, np.int32), np.array(verts, np.int32))
