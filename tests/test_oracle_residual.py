"""Pins the oracle's connectivity, geometry, residual, RK stages and time step
by properties the mathematics fixes: Euler characteristic, free-stream
preservation, global conservation, exact advection of polynomials, the Taylor
polynomials of the RK schemes on a linear operator, the closed-form time step
and the convergence order of the L2 projection."""
import math

import numpy as np
import pytest

from helpers import VORTEX, mesh_tri, np_tri
from oracle import pyoracle as po


def scalar_oracle(nx, p, jitter=0.1, periodic=True, vel=(1.0, 0.5), bc=None):
    vids, xyz, b = mesh_tri(nx, nx, periodic_x=periodic, periodic_y=periodic, jitter=jitter, seed=7,
                            side_tags=(7,) * 4)
    return po.Oracle(2, vids, xyz, p, po.LAW_SCALAR, po.FLUX_UPWIND, vel=vel, farfield=(0.0,), bc=b), vids, xyz


def test_connectivity_torus():
    vids, xyz, bc = mesh_tri(6, 5, jitter=0.1, seed=3)
    o = po.Oracle(2, vids, xyz, 1, po.LAW_SCALAR, po.FLUX_UPWIND, vel=(1, 0))
    f = o.faces()
    E = len(vids)
    assert len(f) == 3 * E // 2
    assert (f[:, 2] >= 0).all()
    V = len(np.unique(vids))
    assert V - len(f) + E == 0  # torus
    # every element side appears exactly once
    sides = set(map(tuple, f[:, :2])) | set(map(tuple, f[:, 2:]))
    assert len(sides) == 3 * E


def test_connectivity_open_square():
    vids, xyz, bc = mesh_tri(4, 3, periodic_x=False, periodic_y=False, side_tags=(1, 2, 3, 4))
    o = po.Oracle(2, vids, xyz, 1, po.LAW_SCALAR, po.FLUX_UPWIND, vel=(1, 0), bc=bc)
    f = o.faces()
    nb = (f[:, 2] < 0).sum()
    assert nb == 2 * (4 + 3)
    assert len(np.unique(vids)) - len(f) + len(vids) == 1  # disc


def test_geometry_area():
    vids, xyz, bc = mesh_tri(8, 8, lx=2.0, ly=3.0, jitter=0.2, seed=5)
    o = po.Oracle(2, vids, xyz, 1, po.LAW_SCALAR, po.FLUX_UPWIND, vel=(1, 0))
    assert abs(o.detj().sum() / 2 - 6.0) < 1e-12


def _constant_state(o, p, val):
    Np = np_tri(p)
    U = np.zeros((o.E, o.nfields, Np))
    phi0 = po.basis(2, 0, (0, 0, 0))[0][0]
    U[:, :, 0] = np.asarray(val) / phi0
    return U.ravel()


@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_scalar_free_stream(p):
    o, _, _ = scalar_oracle(6, p)
    R, _ = o.residual(_constant_state(o, p, [1.3]))
    assert np.abs(R).max() < 1e-13


@pytest.mark.parametrize("flux", [po.FLUX_LF, po.FLUX_ROE])
@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_euler_free_stream(p, flux):
    vids, xyz, bc = mesh_tri(5, 5, jitter=0.15, seed=11)
    q = [1.2, 0.3, -0.7, 3.1]
    o = po.Oracle(2, vids, xyz, p, po.LAW_EULER2D, flux, gamma=1.4)
    R, flag = o.residual(_constant_state(o, p, q))
    assert flag == 0
    assert np.abs(R).max() < 1e-12


@pytest.mark.parametrize("p", [1, 3])
def test_euler_wall_free_stream(p):
    # a fluid at rest next to walls stays at rest; a far-field ghost equal to the
    # state gives no flux jump
    vids, xyz, bc = mesh_tri(5, 4, periodic_x=False, periodic_y=False, jitter=0.1, seed=2,
                             side_tags=(40000, 7, 40000, 7))
    q = [0.9, 0.0, 0.0, 2.0]
    o = po.Oracle(2, vids, xyz, p, po.LAW_EULER2D, po.FLUX_ROE, gamma=1.4, farfield=q, bc=bc)
    R, _ = o.residual(_constant_state(o, p, q))
    assert np.abs(R).max() < 1e-12


@pytest.mark.parametrize("law", ["scalar", "lf", "roe"])
def test_global_conservation(law):
    p = 3
    vids, xyz, bc = mesh_tri(6, 6, lx=10, ly=10, jitter=0.1, seed=9)
    if law == "scalar":
        o = po.Oracle(2, vids, xyz, p, po.LAW_SCALAR, po.FLUX_UPWIND, vel=(1.0, 0.5))
        U = o.project(0)
    else:
        o = po.Oracle(2, vids, xyz, p, po.LAW_EULER2D, po.FLUX_LF if law == "lf" else po.FLUX_ROE, gamma=1.4)
        U = o.project(1, VORTEX)
    R, _ = o.residual(U)
    Np = np_tri(p)
    R = R.reshape(o.E, o.nfields, Np)
    # mode 0 is the constant: sum over cells of its residual is the net boundary flux = 0
    tot = R[:, :, 0].sum(0)
    scale = np.abs(R[:, :, 0]).max()
    assert np.abs(tot).max() < 1e-13 * max(1.0, scale) * o.E


@pytest.mark.parametrize("p", [1, 2, 3])
def test_scalar_polynomial_advection_exact(p):
    # u = polynomial of degree p, continuous: DG residual / |J| in cells away from the
    # boundary equals the L2 projection of -v.grad u (upwind flux of a continuous trace)
    vel = (0.7, -0.4)
    nx = 6
    vids, xyz, bc = mesh_tri(nx, nx, periodic_x=False, periodic_y=False, jitter=0.1, seed=4, side_tags=(7,) * 4)
    o = po.Oracle(2, vids, xyz, p, po.LAW_SCALAR, po.FLUX_UPWIND, vel=vel, farfield=(0.0,), bc=bc)
    c = np.random.default_rng(p).standard_normal(10)

    def u(x, y):
        terms = [1, x, y, x * x, x * y, y * y, x ** 3, x * x * y, x * y * y, y ** 3]
        return sum(ci * t for ci, t in zip(c[: np_tri(p)], terms[: np_tri(p)]))

    def dudt(x, y):
        h = 1e-5
        ux = (u(x + h, y) - u(x - h, y)) / (2 * h)
        uy = (u(x, y + h) - u(x, y - h)) / (2 * h)
        return -(vel[0] * ux + vel[1] * uy)

    Np = np_tri(p)
    # projection by an independent high-order collapsed rule
    from helpers import collapsed_rule
    pts, w = collapsed_rule(p + 3)
    U = np.zeros((o.E, Np))
    Dp = np.zeros((o.E, Np))
    detj = o.detj()
    phis = np.array([po.basis(2, p, (a, b, 0))[0] for a, b in pts])
    for e in range(o.E):
        X = xyz[e]
        xs = X[0, 0] + (X[1, 0] - X[0, 0]) * pts[:, 0] + (X[2, 0] - X[0, 0]) * pts[:, 1]
        ys = X[0, 1] + (X[1, 1] - X[0, 1]) * pts[:, 0] + (X[2, 1] - X[0, 1]) * pts[:, 1]
        U[e] = (phis * (w * u(xs, ys))[:, None]).sum(0)
        Dp[e] = (phis * (w * dudt(xs, ys))[:, None]).sum(0)
    R, _ = o.residual(U.ravel())
    R = R.reshape(o.E, Np) / detj[:, None]
    interior = np.all(bc == 0, axis=1)
    # cells touching the boundary see the far field; skip their neighbours too
    f = o.faces()
    touch = set(f[f[:, 2] < 0][:, 0])
    keep = [e for e in range(o.E) if interior[e] and e not in touch]
    assert len(keep) > 10
    assert np.abs(R[keep] - Dp[keep]).max() < 1e-8 * max(1.0, np.abs(Dp[keep]).max())


@pytest.mark.parametrize("scheme,order", [(po.RK_FE, 1), (po.RK_TVD2, 2), (po.RK_TVD3, 3), (po.RK4, 4)])
def test_rk_schemes_are_taylor_polynomials(scheme, order):
    # for the linear scalar operator L U = R(U)/|J| every scheme here equals
    # sum_{k<=order} (dt L)^k / k! exactly (in exact arithmetic)
    p = 2
    o, _, _ = scalar_oracle(5, p)
    U0 = o.project(0)
    Np = np_tri(p)
    detj = np.repeat(o.detj(), Np)
    dt = 0.01

    def L(U):
        return o.residual(U)[0] / detj

    expect, term = U0.copy(), U0.copy()
    for k in range(1, order + 1):
        term = dt * L(term) / k
        expect = expect + term
    got, _ = o.step(U0, scheme, 0.0, dt)
    assert np.abs(got - expect).max() < 1e-13 * np.abs(U0).max()


def test_dt_closed_form():
    p = 2
    nx = 8
    vel = (1.0, 0.5)
    o, _, _ = scalar_oracle(nx, p, jitter=0.0, vel=vel)
    h = 1.0 / nx
    size = h * (2 - math.sqrt(2))  # 2 x inradius of the right isosceles triangle
    expect = 0.1 * size / ((2 * p + 1) * 1.5)
    assert abs(o.dt(o.project(0), 0.1) - expect) < 1e-15


def test_dt_euler_uses_vertex_values():
    vids, xyz, bc = mesh_tri(4, 4, jitter=0.0)
    o = po.Oracle(2, vids, xyz, 1, po.LAW_EULER2D, po.FLUX_LF, gamma=1.4)
    U = _constant_state(o, 1, [1.0, 0.0, 0.0, 2.5])
    # c = sqrt(1.4 * 1.0 / 1) -> dt = CFL size / (3 c)
    size = 0.25 * (2 - math.sqrt(2))
    assert abs(o.dt(U, 0.5) - 0.5 * size / (3 * math.sqrt(1.4))) < 1e-15


@pytest.mark.parametrize("p", [1, 2, 3])
def test_projection_converges(p):
    errs = []
    for nx in (8, 16):
        vids, xyz, bc = mesh_tri(nx, nx, lx=10, ly=10)
        o = po.Oracle(2, vids, xyz, p, po.LAW_EULER2D, po.FLUX_LF, gamma=1.4)
        U = o.project(1, VORTEX)
        errs.append(o.l2_error(U, 1, VORTEX)[0])
    rate = math.log2(errs[0] / errs[1])
    assert rate > p + 0.5


def test_vortex_exact_is_a_solution():
    # the discrete evolution tracks the exact translated vortex: error after a
    # short RK4 run is at the level of the projection error
    vids, xyz, bc = mesh_tri(12, 12, lx=10, ly=10, jitter=0.1, seed=1)
    o = po.Oracle(2, vids, xyz, 3, po.LAW_EULER2D, po.FLUX_LF, gamma=1.4)
    U = o.project(1, VORTEX)
    e0 = o.l2_error(U, 1, VORTEX, 0.0)[0]
    dt = o.dt(U, 0.1)
    t = 0.0
    for _ in range(10):
        U, _ = o.step(U, po.RK4, t, dt)
        t += dt
    e1 = o.l2_error(U, 1, VORTEX, t)[0]
    assert e1 < 3 * e0


def test_threads_do_not_change_results():
    vids, xyz, bc = mesh_tri(10, 10, lx=10, ly=10, jitter=0.1, seed=1)
    o = po.Oracle(2, vids, xyz, 3, po.LAW_EULER2D, po.FLUX_ROE, gamma=1.4)
    U = o.project(1, VORTEX)
    R1, _ = o.residual(U, nthreads=1)
    R4, _ = o.residual(U, nthreads=4)
    assert np.array_equal(R1, R4)


def test_nonphysical_flag():
    vids, xyz, bc = mesh_tri(4, 4)
    o = po.Oracle(2, vids, xyz, 1, po.LAW_EULER2D, po.FLUX_LF, gamma=1.4)
    U = _constant_state(o, 1, [1.0, 0.0, 0.0, 2.5]).reshape(o.E, 4, 3)
    U[5, 0, 0] = -1.0
    _, flag = o.residual(U.ravel())
    assert flag != 0
