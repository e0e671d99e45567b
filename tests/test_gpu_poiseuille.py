"""The GPU SIMPLE loop on the plane-Poiseuille channel whose discrete fully
developed solution has a closed form (tests/test_oracle_pins.py derives it
from the finite-volume stencil): the converged GPU state must equal that
closed form in the developed section, and the oracle's converged state
everywhere, to 1e-8."""
import numpy as np
import pytest

from anabasis_amd import libpoisson as lp
from anabasis_amd import simple as sm
from oracle import pyoracle as O

from test_oracle_pins import discrete_poiseuille

pytestmark = pytest.mark.gpu


def test_gpu_simple_plane_poiseuille(gpu):
    ny, L, H, W, U, mu = 8, 6.0, 1.0, 0.25, 1.0, 1.0
    nx = int(L / H * ny)
    m = lp.Mesh.hexbox(nx, ny, 1, L, H, W)
    nC, nF = m.nCells, m.nFaces
    ud, Gd = discrete_poiseuille(ny, H, U, mu)
    # patches xmin xmax ymin ymax zmin zmax: inlet, outlet, walls, slip sides
    tU = np.array([1, 0, 1, 1, 0, 0], np.int32)
    tP = np.array([0, 1, 0, 0, 0, 0], np.int32)
    ub = [np.zeros(nF) for _ in range(3)]
    bP = m.bPatch
    sel = np.zeros(nF, bool)
    sel[m.nInternalFaces:] = bP[m.nInternalFaces:] == 1
    j = np.floor(m.xf[sel, 1] / (H / ny)).astype(int)
    ub[0][sel] = ud[j]
    pb = np.zeros(nF)
    par = sm.default_params(rho=1.0, mu=mu, nVelNonOrthCorr=1, nNonOrthCorr=0, uRelax=0.5,
                            pRelax=0.3, corrTol=1e-10, tolMass=1e-11, tolVel=1e-11, velTol=1e-13,
                            velRelTol=0.0, pTol=1e-13, pRelTol=0.0)
    S = sm.Simple(m, par, None, None, inlet_patch="xmin", bc=(tU, tP, *ub, pb))
    u0 = [S.field(k) for k in "uvwp"]
    assert np.allclose(u0[0], ub[0][sel].mean())  # arithmetic inlet mean (main.cu:3302-3322)
    st, conv = S.iterate(3000)
    assert conv and st.step < 3000
    u, v, p = (S.field(k).reshape(ny, nx) for k in "uvp")
    xc = m.array("cc")[:nx, 0]
    cols = (xc >= 0.5 * L) & (xc <= L - 0.5 * H)
    assert np.abs(u[:, cols] - ud[:, None]).max() <= 1e-8 * U
    assert np.abs(v[:, cols]).max() <= 1e-8 * U
    pe = Gd * (xc[cols] - L)
    assert np.abs(p[:, cols] - pe[None, :]).max() <= 1e-8 * np.abs(pe).max()
    # and the oracle, run to its own stop test from the same start
    om = O.OMesh.from_product(m)
    ref = O.simple_run(om, O.simple_params_from(par), tU, tP, *ub, pb, 3000, *u0)
    assert ref["done"] == st.step
    for k in "uvwp":
        a, b = S.field(k), ref[k]
        assert np.abs(a - b).max() <= 1e-8 * max(np.abs(b).max(), U), k
