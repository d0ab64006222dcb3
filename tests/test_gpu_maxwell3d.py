"""GPU: Maxwell on tetrahedra (tiled tet kernel) through the C ABI against the
oracle: residuals P1-P3 with upwind and central fluxes within 1e-12, RK4 steps
within 1e-12, the cavity-mode projection and L2 error, and dt."""
import sys

import numpy as np
import pytest

from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def announce(name):
    print("\n[gpu case] %s" % name, flush=True)
    sys.stdout.flush()


def pair(n, p, flux, use_graph=True):
    import dgpar
    m = dgpar.mesh_tet_structured(n, n, n)
    o = po.Oracle(3, m.vids, m.xyz, p, po.LAW_MAXWELL3D, flux)
    g = dgpar.DGAnalysis(m, dgpar.MaxwellLaw(flux), p, use_graph=use_graph)
    return o, g


def rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.mark.parametrize("flux", [po.FLUX_UPWIND, po.FLUX_CENTRAL])
@pytest.mark.parametrize("p", [1, 2, 3])
def test_tet_residual(flux, p):
    announce("tet residual P%d flux %d" % (p, flux))
    o, g = pair(3, p, flux)
    U = np.random.default_rng(p).standard_normal(o.ndof)
    g.set_state(U)
    assert rel(g.residual(), o.residual(U)[0]) < 1e-12


@pytest.mark.parametrize("use_graph", [True, False])
def test_tet_rk4_steps_and_cavity_error(use_graph):
    announce("tet RK4 graph %d" % use_graph)
    p = 3
    o, g = pair(3, p, po.FLUX_UPWIND, use_graph)
    g.project(3)
    U = o.project(3)
    assert rel(g.get_state(), U) < 1e-13
    dt = o.dt(U, 0.1)
    assert abs(g.computeDt(0.1) - dt) <= 1e-15 * dt
    t = 0.0
    for _ in range(4):
        U, ro = o.step(U, po.RK4, t, dt)
        t += dt
    r = g.advanceInTime(0.0, dt, scheme=po.RK4, nsteps=4)
    assert rel(g.get_state(), U) < 1e-12
    assert abs(r - ro) <= 1e-10 * abs(ro)
    eg = g.l2_error(3, (), t)
    eo = o.l2_error(U, 3, (), t)
    assert np.abs(eg - eo).max() <= 1e-10 * np.abs(eo).max()


def test_tet_rejects_what_it_does_not_build():
    import dgpar
    announce("tet argument errors")
    m = dgpar.mesh_tet_structured(1, 1, 1)
    with pytest.raises(dgpar.DGError):
        dgpar.DGAnalysis(m, dgpar.MaxwellLaw(), 4)  # P4 tets not built
    with pytest.raises(dgpar.DGError):
        dgpar.DGAnalysis(m, dgpar.EulerLaw(), 2)  # Euler on tets not built
    g = dgpar.DGAnalysis(m, dgpar.MaxwellLaw(), 2)
    with pytest.raises(dgpar.DGError):
        g.set_limiter(True)


@pytest.mark.parametrize("shape", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("p", [1, 2, 3])
def test_tet_fp32_form(p, shape):
    """the FP32 form on tetrahedra (FP32 arithmetic, tables and stage copies;
    FP64 step-start state, accumulator and update) against the FP64 oracle:
    the residual within 5e-5 and 3 RK4 steps within 2e-6 (FP32 rounding)"""
    import dgpar
    announce("tet FP32 form P%d shape %d" % (p, shape))
    m = dgpar.mesh_tet_structured(3, 3, 3)
    o = po.Oracle(3, m.vids, m.xyz, p, po.LAW_MAXWELL3D, po.FLUX_UPWIND)
    g = dgpar.DGAnalysis(m, dgpar.MaxwellLaw(po.FLUX_UPWIND), p, precision=dgpar.PREC_FP32)
    g.set_schedule(shape=shape)
    U = np.random.default_rng(p).standard_normal(o.ndof)
    g.set_state(U)
    assert rel(g.residual(), o.residual(U)[0]) < 5e-5
    g.project(3)
    U = o.project(3)
    dt = o.dt(U, 0.1)
    t = 0.0
    for _ in range(3):
        U, _ = o.step(U, po.RK4, t, dt)
        t += dt
    g.advanceInTime(0.0, dt, scheme=po.RK4, nsteps=3)
    assert rel(g.get_state(), U) < 2e-6
