"""GPU: the characteristic far field (tag 30000, EulerLaw.cc:185-278) through
the C ABI against the oracle, which restates the reference's direct flux
formulas and is itself pinned against the reference build:

- residuals P1-P4, subsonic and supersonic free streams, local Lax-Friedrichs,
  Roe and exact fluxes, 4- and 8-wave tiles, within 1e-12;
- RK steps and dt, mixed orders P1-P4, limited steps of a shock tube whose ends
  are characteristic (limiter flags bit for bit), within 1e-12;
- the derivative kernels against finite differences and each other;
- the FP32 form (FP32 arithmetic, FP64 stage accumulation) within 5e-5;
- the matrix-core tile shape and the one-element-per-work-item kernel refuse
  the tag; the other laws read it as a plain far-field tag, as the reference's
  do."""
import sys

import numpy as np
import pytest

from helpers import mesh_tri, np_tri
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

FF = {"subsonic": (1.0, 0.2, -0.1, 2.6), "supersonic": (1.0, 2.0, 0.5, 4.625)}
TAGS = (30000, 30000, 40000, 30000)
FLUX = {"lf": po.FLUX_LF, "roe": po.FLUX_ROE, "exact": po.FLUX_EXACT}


def announce(name):
    print("\n[gpu case] %s" % name, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def pair(regime, flux, p, orders=None, nx=8, **kw):
    import dgpar
    vids, xyz, bc = mesh_tri(nx, nx, periodic_x=False, periodic_y=False, jitter=0.1, seed=6, side_tags=TAGS)
    ff = FF[regime]
    o = po.Oracle(2, vids, xyz, p if orders is None else orders, po.LAW_EULER2D, FLUX[flux], gamma=1.4,
                  farfield=ff, bc=bc)
    g = dgpar.DGAnalysis(dgpar.Mesh(2, vids, xyz, bc, order=orders),
                         dgpar.Law(dgpar.LAW_EULER2D, FLUX[flux], gamma=1.4, farfield=ff),
                         p if orders is None else 2, **kw)
    # small modal noise: at Mach 1.8 a larger one gives P3/P4 quadrature points
    # of negative pressure
    rng = np.random.default_rng(p + 3)
    blocks = []
    for q in (np.full(vids.shape[0], p) if orders is None else orders):
        b = 0.005 * rng.standard_normal((4, np_tri(q)))
        b[:, 0] += np.array(ff) / np.sqrt(2.0)
        blocks.append(b.ravel())
    return g, o, np.concatenate(blocks)


@pytest.mark.parametrize("regime", ["subsonic", "supersonic"])
@pytest.mark.parametrize("flux", ["lf", "roe", "exact"])
@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_residual(regime, flux, p):
    announce("characteristic residual %s %s P%d" % (regime, flux, p))
    g, o, U = pair(regime, flux, p)
    g.set_state(U)
    Ro = o.residual(U)[0]
    assert rel(g.residual(), Ro) < 1e-12
    g.set_schedule(shape=1)  # 8-wave tiles
    g.set_state(U)
    assert rel(g.residual(), Ro) < 1e-12
    import dgpar
    g32, _, _ = pair(regime, flux, p, precision=dgpar.PREC_FP32)
    g32.set_state(U)
    assert rel(g32.residual(), Ro) < 5e-5


@pytest.mark.parametrize("regime,flux,p,scheme", [("subsonic", "roe", 3, po.RK4), ("supersonic", "lf", 2, po.RK_TVD3),
                                                  ("supersonic", "exact", 2, po.RK4)])
def test_steps_and_dt(regime, flux, p, scheme):
    announce("characteristic steps %s %s P%d scheme %d" % (regime, flux, p, scheme))
    g, o, U = pair(regime, flux, p)
    g.set_state(U)
    dt = o.dt(U, 0.1)
    assert abs(g.computeDt(0.1) - dt) <= 1e-13 * dt
    Uo, t = U, 0.0
    for _ in range(3):
        Uo, ro = o.step(Uo, scheme, t, dt)
        t += dt
    rg = g.advanceInTime(0.0, dt, scheme, 3)
    assert rel(g.get_state(), Uo) < 1e-12
    assert abs(rg - ro) <= 1e-10 * abs(ro)


@pytest.mark.parametrize("flux", ["roe", "exact"])
def test_mixed_orders(flux):
    announce("characteristic mixed orders %s" % flux)
    orders = np.random.default_rng(11).integers(1, 5, 2 * 10 * 10).astype(np.int32)
    g, o, U = pair("supersonic", flux, 0, orders=orders, nx=10)
    g.set_state(U)
    assert rel(g.residual(), o.residual(U)[0]) < 1e-12
    dt = o.dt(U, 0.1)
    Uo, t = U, 0.0
    for _ in range(2):
        Uo, _ = o.step(Uo, po.RK_TVD3, t, dt)
        t += dt
    g.advanceInTime(0.0, dt, po.RK_TVD3, 2)
    assert rel(g.get_state(), Uo) < 1e-12


@pytest.mark.parametrize("seed", [None, 7])
def test_limited_shock_tube(seed):
    """a Sod tube with characteristic ends: limiter flags bit for bit, then
    limited SSP-RK3 steps"""
    import dgpar
    announce("characteristic limited shock tube seed %s" % seed)
    ff = (1.0, 0.0, 0.0, 2.5)
    vids, xyz, bc = mesh_tri(16, 4, lx=1.0, ly=0.25, periodic_x=False, periodic_y=False, jitter=0.1, seed=4,
                             side_tags=(40000, 30000, 40000, 30000))
    E = vids.shape[0]
    orders = (np.random.default_rng(seed).integers(1, 5, E).astype(np.int32) if seed is not None
              else np.full(E, 2, np.int32))
    o = po.Oracle(2, vids, xyz, orders, po.LAW_EULER2D, po.FLUX_ROE, gamma=1.4, farfield=ff, bc=bc)
    g = dgpar.DGAnalysis(dgpar.Mesh(2, vids, xyz, bc, order=None if seed is None else orders),
                         dgpar.Law(dgpar.LAW_EULER2D, po.FLUX_ROE, gamma=1.4, farfield=ff), 2)
    U0 = o.project(2, (0.53,))
    g.set_state(U0)
    g.set_limiter(True)
    flags = g.limit()
    U, fo = o.limit(U0)
    assert fo.any() and np.array_equal(flags, fo)
    dt = o.dt(U, 0.1)
    o.set_limiter(True)
    Uo, t = U.copy(), 0.0
    for _ in range(6):
        Uo, _ = o.step(Uo, po.RK_TVD3, t, dt)
        t += dt
    g.set_state(U)
    g.advanceInTime(0.0, dt, scheme=po.RK_TVD3, nsteps=6)
    assert rel(g.get_state(), Uo) < 1e-12


@pytest.mark.parametrize("regime,flux", [("subsonic", "lf"), ("supersonic", "exact")])
def test_derivatives(regime, flux):
    announce("characteristic jvp/vjp %s %s" % (regime, flux))
    g, o, U = pair(regime, flux, 2)
    V = np.random.default_rng(4).standard_normal(U.shape)
    h = 1e-6
    Av = g.residual_jvp(U, V)
    fd = (o.residual(U + h * V)[0] - o.residual(U - h * V)[0]) / (2 * h)
    assert np.abs(Av - fd).max() <= 1e-6 * np.abs(fd).max()
    W = np.random.default_rng(5).standard_normal(U.shape)
    ATw = g.residual_vjp(U, W)
    assert abs(float(np.dot(Av, W)) - float(np.dot(V, ATw))) <= 1e-12 * np.abs(Av).sum() * np.abs(W).max()


def test_fp32_steps():
    import dgpar
    announce("characteristic FP32 RK4 steps")
    g, o, U = pair("supersonic", "roe", 3, precision=dgpar.PREC_FP32)
    g.set_state(U)
    dt = o.dt(U, 0.1)
    Uo, t = U, 0.0
    for _ in range(3):
        Uo, _ = o.step(Uo, po.RK4, t, dt)
        t += dt
    g.advanceInTime(0.0, dt, po.RK4, 3)
    assert rel(g.get_state(), Uo) < 1e-5


def test_refusals_and_other_laws():
    import dgpar
    announce("characteristic refusals; other laws read 30000 as far field")
    g32, _, _ = pair("subsonic", "roe", 2, precision=dgpar.PREC_FP32)
    with pytest.raises(dgpar.DGError):
        g32.set_schedule(shape=2)  # the matrix-core shape is not built for tag 30000
    with pytest.raises(dgpar.DGError, match="tiled"):
        pair("subsonic", "roe", 2, kernel=1)
    vids, xyz, bc = mesh_tri(8, 8, periodic_x=False, periodic_y=False, jitter=0.1, seed=6, side_tags=TAGS)
    o = po.Oracle(2, vids, xyz, 2, po.LAW_MAXWELL2D, po.FLUX_UPWIND, farfield=(0.3, -0.2, 0.5), bc=bc)
    g = dgpar.DGAnalysis(dgpar.Mesh(2, vids, xyz, bc), dgpar.Maxwell2dLaw((0.3, -0.2, 0.5)), 2)
    U = np.random.default_rng(2).standard_normal(o.ndof)
    g.set_state(U)
    assert rel(g.residual(), o.residual(U)[0]) < 1e-12
