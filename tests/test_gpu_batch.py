"""Batched PCG + AMG (an_batch_*): many uneven Poisson MMS systems in one
block-diagonal solve, iterated by one persistent launch or a replayed HIP
graph (the same bits either way)."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from anabasis_amd import _lib as L
from anabasis_amd import libpoisson as lp
from oracle import pyoracle as O

from meshes import bc_arrays

pytestmark = pytest.mark.gpu


def mms_problem(m, seed):
    """Dirichlet Poisson problem with a random smooth manufactured solution
    u = sum_k a_k sin(pi w_k . x + c_k) (f = -lap u), assembled by the
    oracle (orthogonal part)."""
    om = O.OMesh.from_product(m)
    rng = np.random.default_rng(seed)
    a, wv, c = rng.normal(size=3), rng.integers(1, 3, size=(3, 3)), rng.uniform(0, 6.3, 3)
    u = lambda x: sum(a[k] * np.sin(np.pi * x @ wv[k] + c[k]) for k in range(3))
    f = lambda x: sum(a[k] * np.pi ** 2 * (wv[k] @ wv[k]) * np.sin(np.pi * x @ wv[k] + c[k])
                      for k in range(3))
    t, v = bc_arrays(m, dirichlet=tuple(m.patchNames), value=u)
    vals, rhs = O.elliptic_assemble(om, np.ones(om.nFaces), f(om.cc), t, v,
                                    np.zeros((om.nCells, 3)), nonorth=False)
    return om, vals, rhs


@pytest.fixture(scope="module")
def problems(gpu):
    meshes = [lp.Mesh.tetcube(n) for n in (3, 4, 5, 6, 8, 10)] + [
        lp.Mesh.hexbox(12, 9, 7, jitter=0.2, seed=3)]
    out = [mms_problem(m, i) for i, m in enumerate(meshes)]
    om, vals, _ = out[2]
    out.append((om, vals, np.zeros(om.nCells)))  # zero right-hand side
    return out


def test_batch_matches_direct_solves(problems):
    B = lp.BatchPCG([(om.rowOffsets, om.cols) for om, _, _ in problems], relTol=1e-10,
                    absTol=0.0)
    B.set_values([v for _, v, _ in problems])
    xs, infos = B.solve([b for _, _, b in problems])
    rows, secs = B.tuning()
    assert rows in (64, 32, 16, 4, 1) and np.all(secs > 0)
    t = B.tuning_info()
    assert t["rowsPerWave"] == rows and all(v > 0 for v in t["rowSeconds"])
    for (om, vals, b), x, inf in zip(problems, xs, infos):
        if not np.any(b):
            assert np.all(x == 0) and inf.iterations == 0 and inf.converged
            continue
        A = sp.csr_matrix((vals, om.cols, om.rowOffsets))
        xd = spla.spsolve(A.tocsc(), b)
        assert inf.converged and inf.finalRelResNorm <= 1e-10
        assert np.linalg.norm(x - xd) <= 1e-8 * np.linalg.norm(xd)
        r = np.linalg.norm(b - A @ x) / np.linalg.norm(b)
        assert r <= 1.01e-10
    its = [i.iterations for i in infos]
    assert len(set(its)) > 2  # problems stopped independently


def test_batch_graph_replays_equal_eager_and_repeat(problems):
    pats = [(om.rowOffsets, om.cols) for om, _, _ in problems]
    vals = [v for _, v, _ in problems]
    rhs = [b for _, _, b in problems]
    B = lp.BatchPCG(pats, relTol=1e-10, absTol=0.0)
    B.set_values(vals)
    x1, i1 = B.solve(rhs)
    x2, i2 = B.solve(rhs)  # second solve replays the same graph
    B.use_graph(False)
    x3, i3 = B.solve(rhs)
    for a, b, c in zip(x1, x2, x3):
        assert np.array_equal(a, b) and np.array_equal(a, c)
    assert [i.iterations for i in i1] == [i.iterations for i in i3]


def test_batch_mixed_precision(problems):
    pats = [(om.rowOffsets, om.cols) for om, _, _ in problems]
    B = lp.BatchPCG(pats, relTol=1e-10, absTol=0.0, precision=1)
    B.set_values([v for _, v, _ in problems])
    xs, infos = B.solve([b for _, _, b in problems])
    for (om, vals, b), x, inf in zip(problems, xs, infos):
        assert inf.converged
        if np.any(b):
            A = sp.csr_matrix((vals, om.cols, om.rowOffsets))
            assert np.linalg.norm(b - A @ x) <= 1.01e-10 * np.linalg.norm(b)
            assert inf.refinementSteps >= 1


@pytest.mark.parametrize("precision", [0, 1])
def test_batch_persistent_launch_equals_graph_and_eager(problems, precision):
    """One persistent launch per (inner) solve gives the bits of the replayed
    graph and of the eager launches, twice in a row; the tuner timed every
    candidate launch."""
    pats = [(om.rowOffsets, om.cols) for om, _, _ in problems]
    rhs = [b for _, _, b in problems]
    B = lp.BatchPCG(pats, relTol=1e-10, absTol=0.0, precision=precision)
    B.set_values([v for _, v, _ in problems])
    t = B.tuning_info()
    assert t["launch"] in (L.AN_BATCH_GRAPH, L.AN_BATCH_PERSISTENT)
    assert t["launchSeconds"][0] > 0 and t["launchSeconds"][1] > 0
    assert t["persistentBlocksPerCU"] >= 1
    out = {}
    for mode in (L.AN_BATCH_PERSISTENT, L.AN_BATCH_PERSISTENT, L.AN_BATCH_GRAPH,
                 L.AN_BATCH_EAGER):
        B.set_launch(mode)
        xs, infos = B.solve(rhs)
        if mode == L.AN_BATCH_PERSISTENT:
            steps = max(i.refinementSteps for i in infos) if precision == 1 else 1
            assert B.tuning_info()["launchesLastSolve"] == steps
        its = [i.iterations for i in infos]
        if mode in out:
            assert its == out[mode][1]
            assert all(np.array_equal(a, b) for a, b in zip(xs, out[mode][0]))
        out[mode] = (xs, its)
    ref_x, ref_its = out[L.AN_BATCH_EAGER]
    for mode in (L.AN_BATCH_PERSISTENT, L.AN_BATCH_GRAPH):
        assert out[mode][1] == ref_its
        for a, b in zip(out[mode][0], ref_x):
            assert np.array_equal(a, b)
    for (om, vals, b), x in zip(problems, ref_x):
        if np.any(b):
            A = sp.csr_matrix((vals, om.cols, om.rowOffsets))
            assert np.linalg.norm(b - A @ x) <= 1.01e-10 * np.linalg.norm(b)
