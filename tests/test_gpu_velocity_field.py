"""GPU: the scalar law driven by a velocity FieldEvaluator (dg_set_velocity_field,
ScalarLaw.cc:29-97), through the C ABI against the oracle, with the reference's
ScalarPb velocity (2 pi y, -2 pi x) (Lilia.cc:66-82):

- residuals P1-P4 (4- and 8-wave tiles within 1e-12; the FP32 form within 5e-5),
  dt within 1e-14 and SSP-RK3 steps within 1e-12, periodic and walled squares;
- the reference's own mesh files (advection meshes of [-1,1]^2, quadrilaterals
  split) with their 510/610 periodic sides;
- what the velocity field does not serve is refused."""
import os
import sys

import numpy as np
import pytest

from helpers import ROOT, mesh_tri
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def announce(name):
    print("\n[gpu case] %s" % name, flush=True)
    sys.stdout.flush()


def rotating(x, y):
    return (2 * np.pi * y, -2 * np.pi * x)


def rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def pair(vids, xyz, bc, p, precision=0, shape=0, ff=0.4):
    import dgpar
    o = po.Oracle(2, vids, xyz, p, po.LAW_SCALAR, po.FLUX_UPWIND, farfield=(ff,), bc=bc)
    o.set_velocity_field(rotating)
    g = dgpar.DGAnalysis(dgpar.Mesh(2, vids, xyz, bc), dgpar.Law(dgpar.LAW_SCALAR, dgpar.FLUX_UPWIND, farfield=(ff,)),
                         p, precision=precision)
    g.set_velocity_field(rotating)
    g.set_schedule(shape=shape)
    return o, g


@pytest.mark.parametrize("periodic", [True, False])
@pytest.mark.parametrize("shape", [0, 1])
@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_velocity_field_residual_dt_and_steps(p, shape, periodic):
    announce("velocity field P%d shape %d periodic %d" % (p, shape, periodic))
    vids, xyz, bc = mesh_tri(9, 8, x0=-1, y0=-1, lx=2, ly=2, periodic_x=periodic, periodic_y=periodic, jitter=0.1,
                             seed=6, side_tags=(3, 3, 3, 3))
    o, g = pair(vids, xyz, bc, p, shape=shape)
    U = np.random.default_rng(p).standard_normal(o.ndof)
    g.set_state(U)
    assert rel(g.residual(), o.residual(U)[0]) < 1e-12
    dt = o.dt(U, 0.1)
    assert abs(g.computeDt(0.1) - dt) <= 1e-14 * dt
    Uo, t = U, 0.0
    for _ in range(3):
        Uo, ro = o.step(Uo, po.RK_TVD3, t, dt)
        t += dt
    r = g.advanceInTime(0.0, dt, po.RK_TVD3, 3)
    assert rel(g.get_state(), Uo) < 1e-12
    assert abs(r - ro) <= 1e-10 * abs(ro)


@pytest.mark.parametrize("p", [1, 3])
def test_velocity_field_fp32_form(p):
    import dgpar
    announce("velocity field FP32 form P%d" % p)
    vids, xyz, bc = mesh_tri(9, 8, x0=-1, y0=-1, lx=2, ly=2, jitter=0.1, seed=6)
    o, g = pair(vids, xyz, bc, p, precision=dgpar.PREC_FP32)
    U = np.random.default_rng(p).standard_normal(o.ndof)
    g.set_state(U)
    assert rel(g.residual(), o.residual(U)[0]) < 5e-5


@pytest.mark.parametrize("name", ["advection16.msh", "advection256tr.msh"])
def test_scalarpb_on_the_reference_meshes(name):
    import dgpar
    announce("ScalarPb on %s" % name)
    m, info = dgpar.mesh_read_gmsh(os.path.join(ROOT, "tests", "golden", "meshes", name), split_quads=True)
    o, g = pair(m.vids, m.xyz, m.bc, 1)
    U = np.random.default_rng(2).standard_normal(o.ndof)
    g.set_state(U)
    assert rel(g.residual(), o.residual(U)[0]) < 1e-12
    dt = o.dt(U, 0.1)
    Uo = U
    for s in range(4):
        Uo, _ = o.step(Uo, po.RK_TVD2, s * dt, dt)
    g.advanceInTime(0.0, dt, po.RK_TVD2, 4)
    assert rel(g.get_state(), Uo) < 1e-12


def test_velocity_field_refuses_what_it_does_not_serve():
    import dgpar
    announce("velocity field argument errors")
    m = dgpar.mesh_tri_structured(4, 4)
    e = dgpar.DGAnalysis(m, dgpar.EulerLaw(dgpar.FLUX_LF, 1.4), 2)
    with pytest.raises(dgpar.DGError):
        e.set_velocity_field(rotating)
    orders = np.random.default_rng(1).integers(1, 5, m.n_elems).astype(np.int32)
    mx = dgpar.DGAnalysis(dgpar.Mesh(2, m.vids, m.xyz, m.bc, order=orders),
                          dgpar.Law(dgpar.LAW_SCALAR, dgpar.FLUX_UPWIND), 1)
    with pytest.raises(dgpar.DGError):
        mx.set_velocity_field(rotating)
    s = dgpar.DGAnalysis(m, dgpar.Law(dgpar.LAW_SCALAR, dgpar.FLUX_UPWIND), 2)
    s.set_velocity_field(rotating)
    U = np.random.default_rng(1).standard_normal(s.ndof)
    with pytest.raises(dgpar.DGError):
        s.residual_jvp(U, U)
    with pytest.raises(dgpar.DGError):
        s.select_orders((0.05,) * 4, (1e-5,) * 4, apply=True)
