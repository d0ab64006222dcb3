"""GPU residual derivatives and the checkpointed discrete RK4 adjoint, through
the C ABI (dg_residual_jvp / dg_residual_vjp / dg_adjoint_forward /
dg_adjoint_backward):

- tangent-linear product against central finite differences of the oracle's
  residual (h = 1e-6, 1e-6 relative);
- <A v, w> = <v, A^T w> between the GPU tangent and adjoint products (1e-12);
- the gradient of a seeded functional of U(T) against central finite
  differences of the oracle's RK4 steps (h = 1e-6, 1e-6 relative);
- <S, T V> = <g, V> with T the tangent-linear model of the RK4 steps assembled
  from the GPU tangent product (1e-11);
- checkpoint spacing changes nothing (bit-identical gradients), the forward
  state equals plain RK4 steps bit for bit, and the density-misfit objective
  matches its definition."""
import sys

import numpy as np
import pytest

from helpers import FF_OFF, VORTEX, direction, mesh_tri, roughen
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def announce(name):
    print("\n[gpu case] %s" % name, flush=True)
    sys.stdout.flush()


def make(law, flux, p, periodic, nx=6, adj_nd=0):
    import dgpar
    if law == "scalar":
        vids, xyz, bc = mesh_tri(nx, nx, periodic_x=periodic, periodic_y=periodic, jitter=0.1, seed=3,
                                 side_tags=(40000, 7, 40000, 7))
        o = po.Oracle(2, vids, xyz, p, po.LAW_SCALAR, po.FLUX_UPWIND, vel=(1.0, 0.5), farfield=(0.2,), bc=bc)
        lw = dgpar.Law(dgpar.LAW_SCALAR, dgpar.FLUX_UPWIND, vel=(1.0, 0.5), farfield=(0.2,))
        U = o.project(0)
    else:
        vids, xyz, bc = mesh_tri(nx, nx, lx=10, ly=10, periodic_x=periodic, periodic_y=periodic, jitter=0.1, seed=3,
                                 side_tags=(40000, 7, 40000, 7))
        fl = {"lf": po.FLUX_LF, "roe": po.FLUX_ROE, "exact": po.FLUX_EXACT}[flux]
        o = po.Oracle(2, vids, xyz, p, po.LAW_EULER2D, fl, gamma=1.4, farfield=FF_OFF, bc=bc)
        lw = dgpar.Law(dgpar.LAW_EULER2D, fl, gamma=1.4, farfield=FF_OFF)
        U = o.project(1, VORTEX)
    g = dgpar.DGAnalysis(dgpar.Mesh(2, vids, xyz, bc), lw, p, adj_nd=adj_nd)
    return g, o, roughen(U, p)


def rk4_run(o, U, dt, n):
    t = 0.0
    for _ in range(n):
        U, _ = o.step(U, po.RK4, t, dt)
        t += dt
    return U


CASES = [("scalar", None, False), ("euler", "lf", True), ("euler", "roe", True), ("euler", "lf", False),
         ("euler", "roe", False), ("euler", "exact", True), ("euler", "exact", False)]


@pytest.mark.parametrize("law,flux,periodic", CASES)
@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_gpu_jvp_finite_differences_and_dot_product(law, flux, periodic, p):
    announce("jvp/vjp %s %s periodic=%s P%d" % (law, flux, periodic, p))
    g, o, U = make(law, flux, p, periodic)
    h = 1e-6
    for k in range(2):
        V = direction(U, 100 + k)
        Av = g.residual_jvp(U, V)
        fd = (o.residual(U + h * V)[0] - o.residual(U - h * V)[0]) / (2 * h)
        assert np.abs(Av - fd).max() <= 1e-6 * np.abs(fd).max()
        W = np.random.default_rng(300 + k).standard_normal(U.shape)
        ATw = g.residual_vjp(U, W)
        lhs, rhs = float(np.dot(Av, W)), float(np.dot(V, ATw))
        assert abs(lhs - rhs) <= 1e-12 * np.abs(Av).sum() * np.abs(W).max()


@pytest.mark.parametrize("law,flux,periodic,p", [("scalar", None, False, 2), ("euler", "lf", True, 3),
                                                 ("euler", "roe", False, 2), ("euler", "roe", False, 4),
                                                 ("euler", "exact", False, 3)])
def test_gpu_adjoint_gradient_finite_differences(law, flux, periodic, p):
    announce("adjoint FD %s %s periodic=%s P%d" % (law, flux, periodic, p))
    g, o, U0 = make(law, flux, p, periodic)
    dt, n = o.dt(U0, 0.1), 6
    S = np.random.default_rng(11).standard_normal(U0.shape)
    g.set_state(U0)
    g.adjoint_forward(dt, n, 4)
    UT = g.get_state()
    assert np.abs(UT - rk4_run(o, U0, dt, n)).max() <= 1e-12 * np.abs(UT).max()
    grad = g.adjoint_backward(S)
    h = 1e-6
    for k in range(3):
        V = direction(U0, 400 + k)
        fd = (np.dot(S, rk4_run(o, U0 + h * V, dt, n)) - np.dot(S, rk4_run(o, U0 - h * V, dt, n))) / (2 * h)
        assert abs(float(np.dot(grad, V)) - fd) <= 1e-6 * abs(fd)


@pytest.mark.parametrize("law,flux,periodic", CASES)
def test_gpu_adjoint_tangent_linear_identity(law, flux, periodic):
    p = 3
    announce("adjoint TLM identity %s %s periodic=%s P%d" % (law, flux, periodic, p))
    g, o, U0 = make(law, flux, p, periodic)
    dt, n = o.dt(U0, 0.1), 3
    dtj = np.repeat(dt / o.detj(), U0.size // o.detj().size)
    S = np.random.default_rng(12).standard_normal(U0.shape)
    V = direction(U0, 500)
    g.set_state(U0)
    g.adjoint_forward(dt, n, 2)
    grad = g.adjoint_backward(S)
    U, dU = U0, V
    for _ in range(n):
        k, dk = [], []
        Us, dUs = U, dU
        for c in (0.5, 0.5, 1.0, None):
            k.append(dtj * o.residual(Us)[0])
            dk.append(dtj * g.residual_jvp(Us, dUs))
            if c is not None:
                Us, dUs = U + c * k[-1], dU + c * dk[-1]
        U = U + (k[0] + 2 * k[1] + 2 * k[2] + k[3]) / 6
        dU = dU + (dk[0] + 2 * dk[1] + 2 * dk[2] + dk[3]) / 6
    lhs, rhs = float(np.dot(S, dU)), float(np.dot(grad, V))
    assert abs(lhs - rhs) <= 1e-11 * np.abs(S * dU).sum()


def test_gpu_adjoint_checkpoint_spacing_and_objective():
    announce("adjoint checkpoint spacing")
    g, o, U0 = make("euler", "roe", 3, False, nx=8)
    dt, n = o.dt(U0, 0.1), 10
    target = o.project(1, VORTEX)
    grads = []
    for every in (1, 3, 4, 10, 25):
        g.set_state(U0)
        J = g.adjoint_forward(dt, n, every, target)
        grads.append(g.adjoint_backward())
    UT = g.get_state()
    g.set_state(U0)
    g.advanceInTime(0.0, dt, nsteps=n, want_resid=False)
    assert np.array_equal(g.get_state(), UT)
    for gr in grads[1:]:
        assert np.array_equal(gr, grads[0])
    detj = o.detj()
    np_ = U0.size // detj.size // 4
    d = (UT - target).reshape(detj.size, 4, np_)[:, 0, :]
    assert abs(J - 0.5 * np.sum(detj[:, None] * d * d)) <= 1e-12 * J
    seed = np.zeros_like(U0).reshape(detj.size, 4, np_)
    seed[:, 0, :] = detj[:, None] * d
    assert np.array_equal(g.adjoint_backward(seed.ravel()), grads[-1])
    # a forward with fewer steps after a longer one reuses the buffers
    g.set_state(U0)
    g.adjoint_forward(dt, 3, 2)
    g2 = g.adjoint_backward(np.ones_like(U0))
    g.set_state(U0)
    g.adjoint_forward(dt, 3, 3)
    assert np.array_equal(g.adjoint_backward(np.ones_like(U0)), g2)


@pytest.mark.parametrize("p", [1, 3, 4])
def test_gpu_adjoint_directions_per_pass_agree(p):
    announce("adjoint directions per pass P%d" % p)
    grads = []
    for nd in (0, 1, 2, 4):
        g, o, U0 = make("euler", "roe", p, False, adj_nd=nd)
        dt = o.dt(U0, 0.1)
        g.set_state(U0)
        g.adjoint_forward(dt, 3, 2)
        grads.append(g.adjoint_backward(np.random.default_rng(3).standard_normal(U0.shape)))
    for gr in grads[1:]:
        assert np.abs(gr - grads[0]).max() <= 1e-14 * np.abs(grads[0]).max()


def test_gpu_adjoint_rejects_bad_use():
    import dgpar
    announce("adjoint argument errors")
    g, o, U0 = make("euler", "lf", 2, True)
    with pytest.raises(dgpar.DGError):
        g.adjoint_backward(np.ones_like(U0))  # no forward yet
    g.set_state(U0)
    with pytest.raises(dgpar.DGError):
        g.adjoint_forward(0.01, 0, 1)
    g.adjoint_forward(0.01, 2, 1)
    with pytest.raises(dgpar.DGError):
        g.adjoint_backward()  # no target given to the forward
