"""GPU: the scalar law driven by a velocity FieldEvaluator (dg_set_velocity_field,
ScalarLaw.cc:29-97), through the C ABI against the oracle, with the reference's
ScalarPb velocity (2 pi y, -2 pi x) (Lilia.cc:66-82):

- residuals P1-P4 (4- and 8-wave tiles within 1e-12; the FP32 form within 5e-5),
  dt within 1e-14 and SSP-RK3 steps within 1e-12, periodic and walled squares;
- the reference's own mesh files (advection meshes of [-1,1]^2, quadrilaterals
  split) with their 510/610 periodic sides;
- what the velocity field does not serve is refused;
- dg_project_fn (a caller's field projected as DGAnalysis's constructor projects
  the law's far field) within 1e-13 of the oracle's projection, triangles P1-P4
  and tetrahedra (the cavity mode through the callback equals dg_project's)."""
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


@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_project_fn_matches_the_oracle(p):
    """dg_project_fn of ScalarPb's bump (triangles) and of a smooth Euler state
    against the oracle's projection of the same Python field"""
    import dgpar
    announce("project_fn P%d" % p)
    vids, xyz, bc = mesh_tri(10, 10, x0=-1, y0=-1, lx=2, ly=2, jitter=0.1, seed=p)

    def bump(x, y, z, t):
        r = np.hypot(x + 0.5, y)
        return (np.cos(2 * np.pi * r) ** 2 if r <= 0.25 else 0.0,)

    def smooth(x, y, z, t):
        return (1 + 0.2 * np.sin(x + t), 0.3 * np.cos(y), -0.1 * x * y, 2.5 + 0.1 * np.sin(x * y))
    for law, fn, po_law in ((dgpar.Law(dgpar.LAW_SCALAR, dgpar.FLUX_UPWIND), bump, po.LAW_SCALAR),
                            (dgpar.EulerLaw(dgpar.FLUX_LF, 1.4), smooth, po.LAW_EULER2D)):
        g = dgpar.DGAnalysis(dgpar.Mesh(2, vids, xyz, bc), law, p)
        g.project_fn(fn, t=0.25)
        o = po.Oracle(2, vids, xyz, p, po_law, po.FLUX_LF if po_law == po.LAW_EULER2D else po.FLUX_UPWIND,
                      gamma=1.4, bc=bc)
        Uo = o.project_fn(fn, t=0.25)
        assert np.abs(g.get_state() - Uo).max() <= 1e-13 * np.abs(Uo).max()  # as dg_project


def test_project_fn_on_tetrahedra_and_refusals():
    import dgpar
    announce("project_fn tets and argument errors")
    m = dgpar.mesh_tet_structured(2, 2, 2)
    g = dgpar.DGAnalysis(m, dgpar.MaxwellLaw(), 2)
    g.project(dgpar.FIELD_CAVITY, t=0.1)
    ref = g.get_state()
    # the cavity mode through the callback equals the built-in field's projection
    h = dgpar.DGAnalysis(m, dgpar.MaxwellLaw(), 2)
    law = po.make_law(po.LAW_MAXWELL3D, po.FLUX_UPWIND)
    h.project_fn(lambda x, y, z, t: po.exact(3, (x, y, z), t, law=law), t=0.1)
    assert np.abs(h.get_state() - ref).max() <= 1e-13 * np.abs(ref).max()
    with pytest.raises(dgpar.DGError):
        h.project_fn(lambda x, y, z, t: (np.nan,) * 6)
