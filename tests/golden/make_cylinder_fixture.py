"""Writes tests/golden/cylinder_faces.npz from the reference's only mesh
fixture, examples/cylinder/constant/polyMesh/{owner,neighbour,boundary}
(points/faces are absent upstream, .MISSING_LARGE_BLOBS).

The arrays are stored as-is (int32; owner and neighbour delta-encoded for
compression) together with CSR facts computed here by an independent
set-based restatement of build_scalar_pattern (LP/src/scalar_elliptic.cu:
16-58): nnz, a CRC32 of rowOffsets/cols/diagPos/facePP..NN.  Run in a
container that has the reference tree; the GPU tests read only the npz.
"""
import re
import sys
import zlib
from pathlib import Path

import numpy as np

SRC = Path(sys.argv[1] if len(sys.argv) > 1 else
           "/root/reference/examples/cylinder/constant/polyMesh")
OUT = Path(__file__).resolve().parent / "cylinder_faces.npz"


def labels(path):
    txt = path.read_text()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    txt = re.sub(r"//[^\n]*", "", txt)
    m = re.search(r"(\d+)\s*\(", txt)
    n = int(m.group(1))
    body = txt[m.end():txt.index(")", m.end())]
    v = np.array(body.split(), dtype=np.int64)
    assert v.size == n, (path, v.size, n)
    return v.astype(np.int32)


def boundary(path):
    txt = path.read_text()
    out = []
    for name, body in re.findall(r"(\w+)\s*\{([^}]*)\}", txt):
        if name == "FoamFile":
            continue
        nf = int(re.search(r"nFaces\s+(\d+)", body).group(1))
        sf = int(re.search(r"startFace\s+(\d+)", body).group(1))
        out.append((name, sf, nf))
    return out


def reference_pattern(nC, owner, neigh):
    nbrs = [set() for _ in range(nC)]
    for P, N in zip(owner[: neigh.size].tolist(), neigh.tolist()):
        nbrs[P].add(N)
        nbrs[N].add(P)
    ro = [0]
    cols, diag = [], []
    pos = []
    for c in range(nC):
        row = sorted(nbrs[c] | {c})
        pos.append({v: ro[-1] + j for j, v in enumerate(row)})
        diag.append(pos[c][c])
        cols.extend(row)
        ro.append(ro[-1] + len(row))
    PP, PN, NP, NN = [], [], [], []
    for P, N in zip(owner[: neigh.size].tolist(), neigh.tolist()):
        PP.append(pos[P][P])
        PN.append(pos[P][N])
        NP.append(pos[N][P])
        NN.append(pos[N][N])
    arr = [np.array(a, np.int32) for a in (ro, cols, diag, PP, PN, NP, NN)]
    return arr


def crc(arrs):
    c = 0
    for a in arrs:
        c = zlib.crc32(np.ascontiguousarray(a, np.int32).tobytes(), c)
    return c


def main():
    owner = labels(SRC / "owner")
    neigh = labels(SRC / "neighbour")
    patches = boundary(SRC / "boundary")
    nC = int(max(owner.max(), neigh.max())) + 1
    pat = reference_pattern(nC, owner, neigh)
    np.savez_compressed(
        OUT,
        owner_delta=np.diff(owner, prepend=0).astype(np.int32),
        neigh_delta=np.diff(neigh, prepend=0).astype(np.int32),
        patch_names=np.array([p[0] for p in patches]),
        patch_start=np.array([p[1] for p in patches], np.int32),
        patch_nfaces=np.array([p[2] for p in patches], np.int32),
        nCells=np.int64(nC),
        nnz=np.int64(pat[1].size),
        csr_crc=np.int64(crc(pat)),
        row_len_hist=np.bincount(np.diff(pat[0])).astype(np.int64),
    )
    print(OUT, OUT.stat().st_size, "bytes; nC", nC, "nnz", pat[1].size)


if __name__ == "__main__":
    main()
