"""GPU: mixed orders P1-P4 (one persistent launch over order buckets, or one
launch per bucket side by side on side streams: schedule shape 1), the
triangle limiter and the order sensor, through the C ABI, against the oracle.

- mixed-order residual, dt and limited SSP-RK3 / RK4 steps within 1e-12
  (1e-13 for dt) of the oracle;
- limiter flags and sensor orders bit for bit against the oracle on the same
  state; the limited state within 1e-14;
- re-bucketing by new orders carries the state over (raised orders get zero
  modes, lowered ones drop their top modes) and the next residual matches an
  oracle built on the new orders;
- a short shock-tube sequence (limited steps, sensor, re-bucket, limited steps)
  matches the oracle doing the same."""
import sys

import numpy as np
import pytest

from helpers import mesh_tri, np_tri
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

SOD_FF = (1.0, 0.0, 0.0, 2.5, 0.125, 0.0, 0.0, 0.25)
S_HI = (0.05, 0.02, 0.01, 0.005)
S_LO = (5e-5, 2e-5, 1e-5, 5e-6)


def announce(name):
    print("\n[gpu case] %s" % name, flush=True)
    sys.stdout.flush()


def tube(seed, p=2, nx=16, ny=4, flux=po.FLUX_ROE_EF, use_graph=True):
    import dgpar
    vids, xyz, bc = mesh_tri(nx, ny, lx=1.0, ly=0.25, periodic_x=False, periodic_y=False, jitter=0.1, seed=4,
                             side_tags=(40000, 2, 40000, 1))
    E = vids.shape[0]
    orders = (np.random.default_rng(seed).integers(1, 5, E).astype(np.int32) if seed is not None
              else np.full(E, p, np.int32))
    o = po.Oracle(2, vids, xyz, orders, po.LAW_EULER2D, flux, gamma=1.4, farfield=SOD_FF, bc=bc)
    g = dgpar.DGAnalysis(dgpar.Mesh(2, vids, xyz, bc, order=None if seed is None else orders),
                         dgpar.Law(dgpar.LAW_EULER2D, flux, gamma=1.4, farfield=SOD_FF), p, use_graph=use_graph)
    return vids, xyz, bc, orders, o, g


def smooth_state(orders, seed):
    rng = np.random.default_rng(seed)
    out = []
    for p in orders:
        b = 0.01 * rng.standard_normal((4, np_tri(p)))
        b[:, 0] += np.array([0.6, 0.1, -0.05, 1.4]) / np.sqrt(2.0)
        out.append(b.ravel())
    return np.concatenate(out)


def reorder(U, old, new, nf=4):
    out, i = [], 0
    for po_, pn in zip(old, new):
        b = U[i:i + nf * np_tri(po_)].reshape(nf, -1)
        i += nf * np_tri(po_)
        c = np.zeros((nf, np_tri(pn)))
        k = min(b.shape[1], c.shape[1])
        c[:, :k] = b[:, :k]
        out.append(c.ravel())
    return np.concatenate(out)


def rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.mark.parametrize("flux", [po.FLUX_ROE, po.FLUX_LF, po.FLUX_ROE_EF, po.FLUX_EXACT])
@pytest.mark.parametrize("seed", [1, 2])
def test_mixed_residual_and_dt(flux, seed):
    announce("mixed residual flux %d seed %d" % (flux, seed))
    vids, xyz, bc, orders, o, g = tube(seed, flux=flux)
    U = smooth_state(orders, seed)
    g.set_state(U)
    R = o.residual(U)[0]
    assert rel(g.residual(), R) < 1e-12
    g.set_schedule(shape=1)  # one launch per order bucket
    assert rel(g.residual(), R) < 1e-12
    g.set_schedule(shape=0)
    assert abs(g.computeDt(0.1) - o.dt(U, 0.1)) <= 1e-13 * o.dt(U, 0.1)  # vertex sums of up to 15 modes


def test_mixed_scalar_residual():
    import dgpar
    announce("mixed scalar residual")
    vids, xyz, bc = mesh_tri(8, 8, periodic_x=True, periodic_y=False, jitter=0.1, seed=2, side_tags=(7, 7, 7, 7))
    orders = np.random.default_rng(8).integers(1, 5, vids.shape[0]).astype(np.int32)
    o = po.Oracle(2, vids, xyz, orders, po.LAW_SCALAR, po.FLUX_UPWIND, vel=(1.0, 0.5), farfield=(0.2,), bc=bc)
    g = dgpar.DGAnalysis(dgpar.Mesh(2, vids, xyz, bc, order=orders),
                         dgpar.Law(dgpar.LAW_SCALAR, dgpar.FLUX_UPWIND, vel=(1.0, 0.5), farfield=(0.2,)), 1)
    U = np.random.default_rng(3).standard_normal(o.ndof)
    g.set_state(U)
    assert rel(g.residual(), o.residual(U)[0]) < 1e-12
    g.set_schedule(shape=1)
    assert rel(g.residual(), o.residual(U)[0]) < 1e-12


@pytest.mark.parametrize("seed", [None, 5])
@pytest.mark.parametrize("M,nu", [(0.0, 1.0), (5.0, 1.5)])
def test_limiter_flags_bit_exact(seed, M, nu):
    announce("limiter flags seed %s M %g" % (seed, M))
    vids, xyz, bc, orders, o, g = tube(seed, p=3)
    U = o.project(2, (0.53,))
    g.set_state(U)
    g.set_limiter(True, M, nu)
    flags = g.limit()
    Uo, fo = o.limit(U, M, nu)
    assert fo.any()
    assert np.array_equal(flags, fo)
    assert np.array_equal(g.limiter_flags(), fo)
    assert rel(g.get_state(), Uo) < 1e-14


@pytest.mark.parametrize("seed,scheme,use_graph,shape", [(7, po.RK_TVD3, True, 0), (7, po.RK_TVD3, False, 0),
                                                         (None, po.RK_TVD3, True, 0), (9, po.RK4, True, 0),
                                                         (7, po.RK_TVD3, True, 1), (9, po.RK4, False, 1)])
def test_limited_steps(seed, scheme, use_graph, shape):
    announce("limited steps seed %s scheme %d graph %d shape %d" % (seed, scheme, use_graph, shape))
    vids, xyz, bc, orders, o, g = tube(seed, p=2, use_graph=use_graph)
    g.set_schedule(shape=shape)
    U, _ = o.limit(o.project(2, (0.53,)))
    dt = o.dt(U, 0.1)
    o.set_limiter(True)
    Uo, t = U.copy(), 0.0
    for _ in range(6):
        Uo, ro = o.step(Uo, scheme, t, dt)
        t += dt
    g.set_state(U)
    g.set_limiter(True)
    r = g.advanceInTime(0.0, dt, scheme=scheme, nsteps=6)
    assert rel(g.get_state(), Uo) < 1e-12
    assert abs(r - ro) <= 1e-10 * abs(ro)


@pytest.mark.parametrize("p", [2, 3])
def test_uniform_order_limited_steps(p):
    """uniform orders: the tiled stage kernel followed by the limiter"""
    announce("uniform P%d limited steps" % p)
    vids, xyz, bc, orders, o, g = tube(None, p=p)
    U, _ = o.limit(o.project(2, (0.53,)))
    dt = o.dt(U, 0.1)
    o.set_limiter(True)
    Uo, t = U.copy(), 0.0
    for _ in range(5):
        Uo, _ = o.step(Uo, po.RK_TVD3, t, dt)
        t += dt
    g.set_state(U)
    g.set_limiter(True)
    g.advanceInTime(0.0, dt, scheme=po.RK_TVD3, nsteps=5)
    assert rel(g.get_state(), Uo) < 1e-12


@pytest.mark.parametrize("seed", [None, 11])
def test_sensor_orders_and_rebucket(seed):
    announce("sensor seed %s" % seed)
    vids, xyz, bc, orders, o, g = tube(seed, p=3)
    U, _ = o.limit(o.project(2, (0.53,)))
    g.set_state(U)
    want = o.sensor(U, S_HI, S_LO, 1, 4)
    got = g.select_orders(S_HI, S_LO, 1, 4, apply=False)
    assert np.array_equal(got, want)
    assert np.array_equal(g.orders(), orders)
    g.select_orders(S_HI, S_LO, 1, 4, apply=True)
    assert np.array_equal(g.orders(), want)
    V = reorder(U, orders, want)
    assert g.ndof == V.size
    assert np.array_equal(g.get_state(), V)
    o2 = po.Oracle(2, vids, xyz, want, po.LAW_EULER2D, po.FLUX_ROE_EF, gamma=1.4, farfield=SOD_FF, bc=bc)
    W = smooth_state(want, 4)
    g.set_state(W)
    assert rel(g.residual(), o2.residual(W)[0]) < 1e-12


def test_shock_tube_sequence():
    """limited SSP-RK3 steps, sensor + re-bucket, limited steps: GPU and oracle"""
    announce("shock tube sequence")
    vids, xyz, bc, orders, o, g = tube(13, nx=32, ny=8)
    U, _ = o.limit(o.project(2, (0.53,)))
    dt = 0.5 * o.dt(U, 0.1)
    g.set_state(U)
    g.set_limiter(True)
    o.set_limiter(True)
    t = 0.0
    for _ in range(5):
        U, _ = o.step(U, po.RK_TVD3, t, dt)
        t += dt
    g.advanceInTime(0.0, dt, scheme=po.RK_TVD3, nsteps=5)
    assert rel(g.get_state(), U) < 1e-12
    Ug = g.get_state()
    new = g.select_orders(S_HI, S_LO, 1, 4, apply=True)
    assert np.array_equal(new, o.sensor(Ug, S_HI, S_LO, 1, 4))  # same state: same orders
    assert not np.array_equal(new, orders)
    o2 = po.Oracle(2, vids, xyz, new, po.LAW_EULER2D, po.FLUX_ROE_EF, gamma=1.4, farfield=SOD_FF, bc=bc)
    o2.set_limiter(True)
    U2 = reorder(U, orders, new)
    for _ in range(5):
        U2, _ = o2.step(U2, po.RK_TVD3, t, dt)
        t += dt
    g.advanceInTime(0.0, dt, scheme=po.RK_TVD3, nsteps=5)
    assert rel(g.get_state(), U2) < 1e-11


def test_rebucket_in_place_twice_and_back_to_uniform():
    """mixed -> mixed twice (in place on the device), then every element at one
    order (the uniform path, built anew): the state carries over each time and the
    residual matches an oracle built on the new orders"""
    announce("rebucket in place twice, then uniform")
    vids, xyz, bc, orders, o, g = tube(17, nx=24, ny=8)
    cur = orders
    U = smooth_state(cur, 3)
    g.set_state(U)
    for step in range(2):
        # sensor thresholds that raise every element (step 0: top < 2 all always,
        # top > 2 all never) and then lower every element (step 1: top > -all)
        want = np.clip(cur + (1 if step == 0 else -1), 1, 4).astype(np.int32)
        s_hi = (2.0,) * 4 if step == 0 else (-1.0,) * 4
        s_lo = (2.0,) * 4 if step == 0 else (0.0,) * 4
        got = g.select_orders(s_hi, s_lo, 1, 4, apply=True)
        assert np.array_equal(got, want) and np.array_equal(g.orders(), want)
        V = reorder(U, cur, want)
        assert g.ndof == V.size and np.array_equal(g.get_state(), V)
        o2 = po.Oracle(2, vids, xyz, want, po.LAW_EULER2D, po.FLUX_ROE_EF, gamma=1.4, farfield=SOD_FF, bc=bc)
        W = smooth_state(want, 10 + step)
        g.set_state(W)
        assert rel(g.residual(), o2.residual(W)[0]) < 1e-12
        assert abs(g.computeDt(0.1) - o2.dt(W, 0.1)) <= 1e-13 * o2.dt(W, 0.1)
        U, cur = W, want
    # every element down to order 1, one order per sensor pass (the sensor moves
    # an order by one): the last pass leaves a uniform context
    for _ in range(3):
        g.select_orders((-1.0,) * 4, (0.0,) * 4, 1, 4, apply=True)
    ones = np.ones_like(cur)
    assert np.array_equal(g.orders(), ones)
    o3 = po.Oracle(2, vids, xyz, ones, po.LAW_EULER2D, po.FLUX_ROE_EF, gamma=1.4, farfield=SOD_FF, bc=bc)
    W = smooth_state(ones, 20)
    g.set_state(W)
    assert rel(g.residual(), o3.residual(W)[0]) < 1e-12


def test_rebucket_keeps_schedule_and_steps_match_oracle():
    announce("rebucket keeps schedule")
    vids, xyz, bc, orders, o, g = tube(19, nx=32, ny=8)
    U, _ = o.limit(o.project(2, (0.53,)))
    g.set_state(U)
    g.set_schedule(grid=2)
    new = g.select_orders(S_HI, S_LO, 1, 4, apply=True)
    assert g.schedule() == (0, 2)
    o2 = po.Oracle(2, vids, xyz, new, po.LAW_EULER2D, po.FLUX_ROE_EF, gamma=1.4, farfield=SOD_FF, bc=bc)
    V = reorder(U, orders, new)
    dt = 0.5 * o2.dt(V, 0.1)
    Vo, t = V.copy(), 0.0
    for _ in range(3):
        Vo, ro = o2.step(Vo, po.RK_TVD3, t, dt)
        t += dt
    r = g.advanceInTime(0.0, dt, scheme=po.RK_TVD3, nsteps=3)
    assert rel(g.get_state(), Vo) < 1e-12
    assert abs(r - ro) <= 1e-10 * abs(ro)
