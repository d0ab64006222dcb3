"""GPU: meshes read from Gmsh files through the C ABI, against the oracle on the
same descriptor.

- the reference's airfoil2d.msh (2094 triangles, curved-wall tag 20000 on every
  boundary edge): Euler residual (LF, Roe) and RK4 steps within 1e-12;
- the reference's advection256tr.msh (triangles and split quadrilaterals, 510/610
  periodic sides): scalar upwind residual and SSP-RK3 steps within 1e-12;
- a Kuhn tetrahedral mesh written as a Gmsh file: Maxwell residual within 1e-12."""
import os
import sys

import numpy as np
import pytest

from helpers import np_tri, write_msh
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
MESHES = os.path.join(HERE, "golden", "meshes")


def announce(name):
    print("\n[gpu case] %s" % name, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.mark.parametrize("flux", [po.FLUX_LF, po.FLUX_ROE])
def test_airfoil_euler(flux):
    import dgpar
    announce("airfoil2d.msh Euler flux %d" % flux)
    m, _ = dgpar.mesh_read_gmsh(os.path.join(MESHES, "airfoil2d.msh"))
    ff = (1.0, 0.5, 0.1, 2.7)
    p = 2
    o = po.Oracle(2, m.vids, m.xyz, p, po.LAW_EULER2D, flux, gamma=1.4, farfield=ff, bc=m.bc)
    g = dgpar.DGAnalysis(m, dgpar.Law(dgpar.LAW_EULER2D, flux, gamma=1.4, farfield=ff), p)
    rng = np.random.default_rng(3)
    U = 0.01 * rng.standard_normal((m.vids.shape[0], 4, np_tri(p)))
    U[:, :, 0] += np.array(ff) / np.sqrt(2.0)
    U = U.ravel()
    g.set_state(U)
    assert rel(g.residual(), o.residual(U)[0]) < 1e-12
    dt = o.dt(U, 0.1)
    Uo, t = U.copy(), 0.0
    for _ in range(3):
        Uo, ro = o.step(Uo, po.RK4, t, dt)
        t += dt
    r = g.advanceInTime(0.0, dt, scheme=po.RK4, nsteps=3)
    assert rel(g.get_state(), Uo) < 1e-12
    assert abs(r - ro) <= 1e-10 * abs(ro)


def test_advection_torus_scalar():
    import dgpar
    announce("advection256tr.msh scalar")
    m, _ = dgpar.mesh_read_gmsh(os.path.join(MESHES, "advection256tr.msh"), split_quads=True)
    p = 2
    o = po.Oracle(2, m.vids, m.xyz, p, po.LAW_SCALAR, po.FLUX_UPWIND, vel=(1.0, 0.5), farfield=(0.0,), bc=m.bc)
    g = dgpar.DGAnalysis(m, dgpar.Law(dgpar.LAW_SCALAR, dgpar.FLUX_UPWIND, vel=(1.0, 0.5), farfield=(0.0,)), p)
    U = np.random.default_rng(5).standard_normal(o.ndof)
    g.set_state(U)
    assert rel(g.residual(), o.residual(U)[0]) < 1e-12
    dt = o.dt(U, 0.1)
    Uo, t = U.copy(), 0.0
    for _ in range(3):
        Uo, _ = o.step(Uo, po.RK_TVD3, t, dt)
        t += dt
    g.advanceInTime(0.0, dt, scheme=po.RK_TVD3, nsteps=3)
    assert rel(g.get_state(), Uo) < 1e-12


def test_kuhn_msh_maxwell(tmp_path):
    import dgpar
    announce("Kuhn .msh Maxwell")
    k = dgpar.mesh_tet_structured(3, 3, 3)
    keys, nodes, recs = {}, [], []
    for e in range(k.vids.shape[0]):
        ids = []
        for v in range(4):
            key = tuple(np.round(k.xyz[e, v], 12))
            if key not in keys:
                keys[key] = len(nodes) + 1
                nodes.append(key)
            ids.append(keys[key])
        recs.append((4, 100, ids))
    path = tmp_path / "kuhn.msh"
    write_msh(path, nodes, recs)
    m, info = dgpar.mesh_read_gmsh(path)
    assert info["dim"] == 3
    p = 2
    o = po.Oracle(3, m.vids, m.xyz, p, po.LAW_MAXWELL3D, po.FLUX_UPWIND)
    g = dgpar.DGAnalysis(m, dgpar.MaxwellLaw(po.FLUX_UPWIND), p)
    U = np.random.default_rng(7).standard_normal(o.ndof)
    g.set_state(U)
    assert rel(g.residual(), o.residual(U)[0]) < 1e-12
