"""The discrete RK4 adjoint as a libtorch autograd Function
(dg-parallel_amd/autograd/dg_torch.cpp over the C ABI): torch's gradcheck on a
linear law (central differences are then exact up to rounding), the gradient
of a loss through torch.autograd against the library's adjoint called directly
and against central differences of the oracle's RK4 steps, and a backward
that follows another forward on the same context (it re-runs its own)."""
import os
import sys

import numpy as np
import pytest

from helpers import FF_OFF, ROOT, VORTEX, direction, mesh_tri, roughen
from oracle import pyoracle as po

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not os.path.exists(os.path.join(ROOT, "dg-parallel_amd", "autograd", "dgpar_torch.so")),
                                 reason="autograd extension not built")]


def announce(name):
    print("\n[gpu case] %s" % name, flush=True)
    sys.stdout.flush()


def test_gradcheck_scalar_law():
    import torch  # host tensors only: torch never starts its own GPU runtime here
    import dgpar
    import dgpar_autograd
    announce("torch gradcheck scalar")
    vids, xyz, bc = mesh_tri(3, 3, periodic_x=True, periodic_y=True, jitter=0.1, seed=3)
    g = dgpar.DGAnalysis(dgpar.Mesh(2, vids, xyz, bc), dgpar.ScalarLaw((1.0, 0.5, 0.0)), 1)
    U0 = torch.tensor(np.random.default_rng(1).standard_normal(g.ndof), dtype=torch.float64, requires_grad=True)
    assert torch.autograd.gradcheck(lambda u: dgpar_autograd.rk4(u, g, 0.01, 3, 2), (U0,), eps=1e-6, atol=1e-8,
                                    rtol=1e-6)


def test_loss_gradient_matches_adjoint_and_finite_differences():
    import torch
    import dgpar
    import dgpar_autograd
    announce("torch loss gradient euler")
    p, n = 2, 5
    vids, xyz, bc = mesh_tri(6, 6, lx=10, ly=10, periodic_x=False, periodic_y=False, jitter=0.1, seed=3,
                             side_tags=(40000, 7, 40000, 7))
    o = po.Oracle(2, vids, xyz, p, po.LAW_EULER2D, po.FLUX_ROE, gamma=1.4, farfield=FF_OFF, bc=bc)
    U0 = roughen(o.project(1, VORTEX), p)
    dt = o.dt(U0, 0.1)
    g = dgpar.DGAnalysis(dgpar.Mesh(2, vids, xyz, bc), dgpar.Law(dgpar.LAW_EULER2D, dgpar.FLUX_ROE, gamma=1.4,
                                                                 farfield=FF_OFF), p)
    W = np.random.default_rng(4).standard_normal(U0.shape)
    u = torch.tensor(U0, dtype=torch.float64, requires_grad=True)
    UT = dgpar_autograd.rk4(u, g, dt, n, 2)
    loss = (torch.tensor(W) * UT).sum() + 0.5 * (UT ** 2).sum()
    loss.backward()
    grad = u.grad.numpy()
    # the same seed through the C ABI directly
    seed = W + UT.detach().numpy()
    g.set_state(U0)
    g.adjoint_forward(dt, n, 3)
    assert np.array_equal(g.adjoint_backward(seed), grad) or \
        np.abs(g.adjoint_backward(seed) - grad).max() <= 1e-14 * np.abs(grad).max()

    def run(U):
        t = 0.0
        for _ in range(n):
            U, _ = o.step(U, po.RK4, t, dt)
            t += dt
        return float(np.dot(W, U) + 0.5 * np.dot(U, U))

    h = 1e-6
    for k in range(2):
        V = direction(U0, 700 + k)
        fd = (run(U0 + h * V) - run(U0 - h * V)) / (2 * h)
        assert abs(float(np.dot(grad, V)) - fd) <= 1e-6 * abs(fd)


def test_backward_after_another_forward_on_the_same_context():
    import torch
    import dgpar
    import dgpar_autograd
    announce("torch interleaved forwards")
    vids, xyz, bc = mesh_tri(4, 4, periodic_x=True, periodic_y=True, jitter=0.1, seed=3)
    g = dgpar.DGAnalysis(dgpar.Mesh(2, vids, xyz, bc), dgpar.ScalarLaw((1.0, 0.5, 0.0)), 2)
    rng = np.random.default_rng(2)
    a = torch.tensor(rng.standard_normal(g.ndof), dtype=torch.float64, requires_grad=True)
    b = torch.tensor(rng.standard_normal(g.ndof), dtype=torch.float64, requires_grad=True)
    ya = dgpar_autograd.rk4(a, g, 0.01, 4, 2)
    yb = dgpar_autograd.rk4(b, g, 0.01, 4, 2)   # overwrites the context's checkpoints
    (ya ** 2).sum().backward()
    (yb ** 2).sum().backward()
    a2 = a.detach().clone().requires_grad_(True)
    (dgpar_autograd.rk4(a2, g, 0.01, 4, 2) ** 2).sum().backward()
    assert torch.equal(a.grad, a2.grad)
