"""Residual derivatives and the discrete RK4 adjoint, run through the CPU
emulation of their HIP kernels (tests/cpp/emu_stage.cc):

- tangent-linear products against central finite differences of the oracle's
  residual (h = 1e-6, 1e-6 relative);
- the adjoint product against the tangent-linear one by the dot-product
  identity <A v, w> = <v, A^T w> (1e-12 relative);
- the gradient of <s, U(T)> after several RK4 steps against central finite
  differences of the oracle's RK4 steps along seeded directions (1e-6).

The finite differences are taken on the oracle, so a wrong flux Jacobian, a
dropped face term or a wrong reverse-stage coefficient fails here."""
import ctypes

import numpy as np
import pytest

from oracle import pyoracle as po
from helpers import FF_OFF, direction, roughen
from test_kernel_emulation import LIBEMU, case, pytestmark  # noqa: F401

PD = ctypes.POINTER(ctypes.c_double)
if LIBEMU is not None:
    import dgpar

    p_md, p_ld = ctypes.POINTER(dgpar.MeshDesc), ctypes.POINTER(dgpar.LawDesc)
    LIBEMU.emu_tri_derivative.argtypes = [p_md, p_ld, ctypes.c_int, ctypes.c_int, PD, PD, PD]
    LIBEMU.emu_tri_adjoint_rk4.argtypes = [p_md, p_ld, ctypes.c_int, ctypes.c_double, ctypes.c_int, PD, PD, PD,
                                           PD]

CASES = [("scalar", None, True), ("scalar", None, False), ("euler", "lf", True), ("euler", "roe", True),
         ("euler", "lf", False), ("euler", "roe", False)]


def emu_derivative(mesh, lw, p, U, V, jvp, nd=0):
    md, ld = mesh.desc(), lw.desc()
    U = np.ascontiguousarray(U)
    V = np.ascontiguousarray(V)
    W = np.zeros_like(U)
    assert LIBEMU.emu_tri_derivative(ctypes.byref(md), ctypes.byref(ld), p, 1 if jvp else -nd, U.ctypes.data_as(PD),
                                     V.ctypes.data_as(PD), W.ctypes.data_as(PD)) == 0
    return W


def emu_adjoint(mesh, lw, p, dt, n, U0, seed):
    md, ld = mesh.desc(), lw.desc()
    U0 = np.ascontiguousarray(U0)
    seed = np.ascontiguousarray(seed)
    UT, g = np.zeros_like(U0), np.zeros_like(U0)
    assert LIBEMU.emu_tri_adjoint_rk4(ctypes.byref(md), ctypes.byref(ld), p, dt, n, U0.ctypes.data_as(PD),
                                      seed.ctypes.data_as(PD), UT.ctypes.data_as(PD), g.ctypes.data_as(PD)) == 0
    return UT, g


@pytest.mark.parametrize("law,flux,periodic", CASES)
@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_jvp_matches_finite_differences(law, flux, periodic, p):
    mesh, lw, o, U = case(law, flux, p, periodic, nx=6, ff=FF_OFF)
    U = roughen(U, p)
    h = 1e-6
    for k in range(2):
        V = direction(U, 100 + k)
        Av = emu_derivative(mesh, lw, p, U, V, jvp=True)
        fd = (o.residual(U + h * V)[0] - o.residual(U - h * V)[0]) / (2 * h)
        assert np.abs(Av - fd).max() <= 1e-6 * np.abs(fd).max(), (k, np.abs(Av - fd).max() / np.abs(fd).max())


@pytest.mark.parametrize("law,flux,periodic", CASES)
@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_vjp_dot_product_identity(law, flux, periodic, p):
    mesh, lw, o, U = case(law, flux, p, periodic, nx=6, ff=FF_OFF)
    for k in range(3):
        V = direction(U, 200 + k)
        Wt = np.random.default_rng(300 + k).standard_normal(U.shape)
        Av = emu_derivative(mesh, lw, p, U, V, jvp=True)
        ATw = emu_derivative(mesh, lw, p, U, Wt, jvp=False)
        lhs, rhs = float(np.dot(Av, Wt)), float(np.dot(V, ATw))
        scale = np.abs(Av).sum() * np.abs(Wt).max() + 1e-300
        assert abs(lhs - rhs) <= 1e-12 * scale, (lhs, rhs)


@pytest.mark.parametrize("flux,periodic", [("lf", False), ("roe", False), ("roe", True)])
@pytest.mark.parametrize("p", [1, 3, 4])
def test_vjp_directions_per_pass_agree(flux, periodic, p):
    """the adjoint kernel builds each flux Jacobian in NF/ND passes; every split
    gives the same product (the dual arithmetic per direction is unchanged)"""
    mesh, lw, o, U = case("euler", flux, p, periodic, nx=6, ff=FF_OFF)
    U = roughen(U, p)
    Z = np.random.default_rng(9).standard_normal(U.shape)
    ref = emu_derivative(mesh, lw, p, U, Z, jvp=False)
    for nd in (1, 2):
        assert np.abs(emu_derivative(mesh, lw, p, U, Z, jvp=False, nd=nd) - ref).max() <= 1e-14 * np.abs(ref).max()


def test_vjp_columns_by_brute_force():
    """A^T e_i against the i-th row of A assembled column by column from the
    tangent-linear product, on a tiny mesh (every entry of the transpose)"""
    mesh, lw, o, U = case("euler", "roe", 1, False, nx=2, ff=FF_OFF)
    n = U.size
    A = np.zeros((n, n))
    for j in range(n):
        e = np.zeros(n)
        e[j] = 1.0
        A[:, j] = emu_derivative(mesh, lw, 1, U, e, jvp=True)
    Z = np.random.default_rng(5).standard_normal(n)
    ATz = emu_derivative(mesh, lw, 1, U, Z, jvp=False)
    assert np.abs(ATz - A.T @ Z).max() <= 1e-12 * np.abs(A.T @ Z).max()


def rk4_run(o, U, dt, nsteps):
    t = 0.0
    for _ in range(nsteps):
        U, _ = o.step(U, po.RK4, t, dt)
        t += dt
    return U


# Lax-Friedrichs on a bounded domain is left to the tangent-linear identity below:
# its wave-speed maximum switches sides at some face point within 1e-6 of these
# 4-step trajectories (one-sided differences disagree at 4e-6 and meet as h -> 0)
@pytest.mark.parametrize("law,flux,periodic,p", [("scalar", None, False, 2), ("euler", "lf", True, 3),
                                                 ("euler", "roe", False, 2), ("euler", "roe", False, 4),
                                                 ("euler", "lf", True, 4)])
def test_rk4_adjoint_gradient(law, flux, periodic, p):
    mesh, lw, o, U0 = case(law, flux, p, periodic, nx=6, ff=FF_OFF)
    U0 = roughen(U0, p)
    dt, nsteps = o.dt(U0, 0.1), 4
    S = np.random.default_rng(11).standard_normal(U0.shape)
    UT, g = emu_adjoint(mesh, lw, p, dt, nsteps, U0, S)
    assert np.abs(UT - rk4_run(o, U0, dt, nsteps)).max() <= 1e-12 * np.abs(UT).max()
    h = 1e-6
    for k in range(3):
        V = direction(U0, 400 + k)
        fd = (np.dot(S, rk4_run(o, U0 + h * V, dt, nsteps)) - np.dot(S, rk4_run(o, U0 - h * V, dt, nsteps))) / (2 * h)
        ad = float(np.dot(g, V))
        assert abs(ad - fd) <= 1e-6 * abs(fd), (k, ad, fd)


@pytest.mark.parametrize("law,flux,periodic", CASES)
@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_rk4_adjoint_tangent_linear_identity(law, flux, periodic, p):
    """<S, T V> = <g, V> with T the tangent-linear model of 3 RK4 steps, built
    here stage by stage from the emulated residual tangent (dk = dt/|J| A(U_s) dU_s)
    on the oracle's stage states: holds to rounding whatever kinks the flux has"""
    mesh, lw, o, U0 = case(law, flux, p, periodic, nx=6, ff=FF_OFF)
    U0 = roughen(U0, p)
    dt, nsteps = o.dt(U0, 0.1), 3
    dtj = np.repeat(dt / o.detj(), U0.size // o.detj().size)
    S = np.random.default_rng(12).standard_normal(U0.shape)
    V = direction(U0, 500)
    _, g = emu_adjoint(mesh, lw, p, dt, nsteps, U0, S)
    U, dU = U0, V
    for _ in range(nsteps):
        k, dk = [], []
        Us, dUs = U, dU
        for s, c in enumerate((0.5, 0.5, 1.0, None)):
            k.append(dtj * o.residual(Us)[0])
            dk.append(dtj * emu_derivative(mesh, lw, p, Us, dUs, jvp=True))
            if c is not None:
                Us, dUs = U + c * k[-1], dU + c * dk[-1]
        U = U + (k[0] + 2 * k[1] + 2 * k[2] + k[3]) / 6
        dU = dU + (dk[0] + 2 * dk[1] + 2 * dk[2] + dk[3]) / 6
    lhs, rhs = float(np.dot(S, dU)), float(np.dot(g, V))
    assert abs(lhs - rhs) <= 1e-11 * np.abs(S * dU).sum(), (lhs, rhs)
