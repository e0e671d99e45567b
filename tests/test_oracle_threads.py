"""The oracle's parallel loops (OpenMP) give the same bits on any thread
count: a PCG+AMG solve, an LSQ gradient and three SIMPLE outer iterations
computed here (all threads) equal the same computation in a child process
restricted to one thread."""
import os
import subprocess
import sys
from pathlib import Path

import numpy as np

from anabasis_amd.libpoisson import Mesh
from oracle import pyoracle as O

from meshes import bc_arrays

TESTS = Path(__file__).resolve().parent
ROOT = TESTS.parent


def compute(out_path):
    m = Mesh.hexbox(14, 12, 10, 1.4, 1.2, 1.0, jitter=0.2, seed=3)
    om = O.OMesh.from_product(m)
    t, v = bc_arrays(m, dirichlet=("xmin",))
    vals, _ = O.elliptic_assemble(om, np.ones(om.nFaces), np.zeros(om.nCells), t, v,
                                  np.zeros((om.nCells, 3)))
    b = np.random.default_rng(5).standard_normal(om.nCells)
    x, it, rr, _ = O.pcg_amg(om.rowOffsets, om.cols, vals, b, relTol=1e-10, absTol=0.0,
                             coarseSize=32)
    phi = np.sin(om.cc[:, 0]) * np.cos(om.cc[:, 1]) + om.cc[:, 2]
    g = O.lsq_gradient(om, phi, t, v, mode="gsv1")
    nF, nC = om.nFaces, om.nCells
    tU = np.zeros(nF - om.nInternalFaces, np.int32) + 1
    tP = np.zeros(nF - om.nInternalFaces, np.int32)
    tP[om.bPatch[om.nInternalFaces:] == 2] = 1  # xmax: fixed pressure
    tU[om.bPatch[om.nInternalFaces:] == 2] = 0
    ub = [np.zeros(nF) for _ in range(3)]
    ub[0][om.bPatch == 1] = 0.3  # xmin inflow
    par = O.SimpleParams(1.0, 0.1, 1, 2, 0.5, 0.3, 1e-6, 1e-12, 1e-12, 1000, 1e-10, 0.0, 2000,
                         1e-10, 0.0, 1, O.amg_options(coarseSize=32))
    r = O.simple_run(om, par, tU, tP, *ub, np.zeros(nF), 3, np.full(nC, 0.3), np.zeros(nC),
                     np.zeros(nC), np.zeros(nC))
    np.savez(out_path, x=x, it=it, g=g, u=r["u"], p=r["p"], phi=r["phi"])


def test_oracle_results_independent_of_thread_count(tmp_path):
    many = tmp_path / "many.npz"
    one = tmp_path / "one.npz"
    compute(many)
    env = dict(os.environ, OMP_NUM_THREADS="1")
    code = ("import sys; sys.path[:0] = [%r, %r]; import test_oracle_threads as t; "
            "t.compute(%r)" % (str(ROOT), str(TESTS), str(one)))
    subprocess.run([sys.executable, "-c", code], env=env, check=True, cwd=str(ROOT))
    a, b = np.load(many), np.load(one)
    for k in a.files:
        assert np.array_equal(a[k], b[k]), k
