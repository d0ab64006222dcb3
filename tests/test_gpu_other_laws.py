"""GPU: Maxwell2d and Burgers2D (MaxwellLaw.cc:21-200, Burgers2D.cc:22-92)
through the C ABI against the oracle (itself pinned against the reference build):
residuals P1-P4 on periodic meshes and on meshes with walls and far-field sides
(the tiled kernel and the one-element-per-work-item kernel, 1e-12; the FP32 form,
5e-5), RK steps and dt (1e-12), mixed orders P1-P4 in the persistent launch,
and the derivative kernels against finite differences."""
import sys

import numpy as np
import pytest

from helpers import mesh_tri
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu
FF = {"maxwell2d": (0.3, -0.2, 0.5), "burgers": (0.4,)}


def announce(name):
    print("\n[gpu case] %s" % name, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def pair(law, p, periodic=True, nx=8, orders=None, **kw):
    import dgpar
    vids, xyz, bc = mesh_tri(nx, nx, periodic_x=periodic, periodic_y=periodic, jitter=0.1, seed=5,
                             side_tags=(40000, 7, 40000, 7))
    lo = po.LAW_MAXWELL2D if law == "maxwell2d" else po.LAW_BURGERS2D
    fl = po.FLUX_UPWIND if law == "maxwell2d" else po.FLUX_CENTRAL
    o = po.Oracle(2, vids, xyz, p if orders is None else orders, lo, fl, farfield=FF[law], bc=bc)
    lw = dgpar.Maxwell2dLaw(FF[law]) if law == "maxwell2d" else dgpar.Burgers2DLaw(FF[law])
    g = dgpar.DGAnalysis(dgpar.Mesh(2, vids, xyz, bc, order=orders), lw, p if orders is None else 2, **kw)
    # Burgers2D's central flux has no dissipation and its dt uses wave speed 1:
    # a small state keeps the steps stable
    U = (0.5 if law == "maxwell2d" else 0.1) * np.random.default_rng(p + 1).standard_normal(o.ndof)
    return g, o, U


@pytest.mark.parametrize("law", ["maxwell2d", "burgers"])
@pytest.mark.parametrize("periodic", [True, False])
@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_residual(law, periodic, p):
    import dgpar
    announce("%s residual P%d periodic=%s" % (law, p, periodic))
    g, o, U = pair(law, p, periodic)
    g.set_state(U)
    Ro = o.residual(U)[0]
    assert rel(g.residual(), Ro) < 1e-12
    g8, _, _ = pair(law, p, periodic)
    g8.set_schedule(shape=1)  # 8-wave tiles
    g8.set_state(U)
    assert rel(g8.residual(), Ro) < 1e-12
    g32, _, _ = pair(law, p, periodic, precision=dgpar.PREC_FP32)
    g32.set_state(U)
    assert rel(g32.residual(), Ro) < 5e-5


def test_element_kernel_refused():
    """the one-element-per-work-item kernel runs the scalar and Euler laws only"""
    import dgpar
    announce("element kernel refused for the other laws")
    for law in ("maxwell2d", "burgers"):
        with pytest.raises(dgpar.DGError, match="one-element-per-work-item"):
            pair(law, 2, kernel=1)


@pytest.mark.parametrize("law,p,scheme", [("maxwell2d", 3, po.RK4), ("maxwell2d", 2, po.RK_TVD3),
                                          ("burgers", 2, po.RK_TVD3), ("burgers", 4, po.RK4)])
def test_steps_and_dt(law, p, scheme):
    announce("%s steps P%d scheme %d" % (law, p, scheme))
    g, o, U = pair(law, p, periodic=False)
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


@pytest.mark.parametrize("law", ["maxwell2d", "burgers"])
def test_mixed_orders(law):
    announce("%s mixed orders" % law)
    orders = np.random.default_rng(9).integers(1, 5, 2 * 12 * 12).astype(np.int32)
    g, o, U = pair(law, 0, periodic=False, nx=12, orders=orders)
    g.set_state(U)
    assert rel(g.residual(), o.residual(U)[0]) < 1e-12
    dt = o.dt(U, 0.1)
    Uo, t = U, 0.0
    for _ in range(2):
        Uo, _ = o.step(Uo, po.RK_TVD3, t, dt)
        t += dt
    g.advanceInTime(0.0, dt, po.RK_TVD3, 2)
    assert rel(g.get_state(), Uo) < 1e-12


@pytest.mark.parametrize("law", ["maxwell2d", "burgers"])
def test_derivatives(law):
    announce("%s jvp/vjp" % law)
    g, o, U = pair(law, 3, periodic=False)
    V = np.random.default_rng(4).standard_normal(U.shape)
    h = 1e-6
    Av = g.residual_jvp(U, V)
    fd = (o.residual(U + h * V)[0] - o.residual(U - h * V)[0]) / (2 * h)
    assert np.abs(Av - fd).max() <= 1e-6 * np.abs(fd).max()
    W = np.random.default_rng(5).standard_normal(U.shape)
    ATw = g.residual_vjp(U, W)
    assert abs(float(np.dot(Av, W)) - float(np.dot(V, ATw))) <= 1e-12 * np.abs(Av).sum() * np.abs(W).max()


def test_limiter_refused():
    import dgpar
    announce("limiter refused for Maxwell2d")
    g, _, _ = pair("maxwell2d", 2)
    with pytest.raises(dgpar.DGError, match="limiter"):
        g.set_limiter(True)


def test_pos_views(tmp_path):
    """the .pos views of both laws carry the sampled fields (Maxwell2d: Ez^2,
    Hx, Hy, Ez, MaxwellLaw.cc:256-330; Burgers2D: the function, Burgers2D.cc:121-200)"""
    announce("pos views of the other laws")
    g, o, U = pair("maxwell2d", 2)
    g.set_state(U)
    g.export_pos(str(tmp_path / "m"), 1, nsub=1)
    S = g.sample_lattice(1)  # the corners, one sub-triangle per element
    for name, val in (("Intensity_RefCond", S[:, 2] ** 2), ("HxField_RefCond", S[:, 0]),
                      ("HyField_RefCond", S[:, 1]), ("EField_RefCond", S[:, 2])):
        lines = open(tmp_path / ("m-sample1%s.pos" % name)).read().splitlines()[1:-1]
        got = np.array([[float(x) for x in ln.split("{")[1].rstrip("};").split(",")] for ln in lines])
        assert got.shape == val.shape and np.abs(got - val).max() <= 1e-8 * max(1.0, np.abs(val).max())
    b, _, Ub = pair("burgers", 2)
    b.set_state(Ub)
    b.export_pos(str(tmp_path / "b"), 0, nsub=1)
    lines = open(tmp_path / "b-sample0function.pos").read().splitlines()[1:-1]
    got = np.array([[float(x) for x in ln.split("{")[1].rstrip("};").split(",")] for ln in lines])
    assert np.abs(got - b.sample_lattice(1)[:, 0]).max() <= 1e-8
