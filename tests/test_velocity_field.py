"""The scalar law driven by a velocity FieldEvaluator (ScalarLaw.cc:29-97: Fi,
riemannSolver, boundaryFlux and maximumEigenValue evaluate velocityField at
their point), with the reference's ScalarPb velocity (2 pi y, -2 pi x)
(VelocityOne, Lilia.cc:66-82):

- the oracle's restatement against the reference's own ScalarLaw with that
  FieldEvaluator (oracle/_ref, REF_VEL=rotate): residuals P1-P4 on jittered
  origin-centred periodic squares (where the reference's reflective
  periodicity is translation and the rotation's normal velocity agrees on the
  two images of a periodic face) and on walled/far-field squares, and RK4 steps;
- the oracle pinned by the mathematics: a field that is constant reproduces the
  constant-velocity law bit for bit; the divergence-free rotation conserves the
  total on a periodic origin-centred square;
- the tiled kernel's own phase functions (CPU emulation) against the oracle."""
import ctypes
import math
import os

import numpy as np
import pytest

import dgpar
import ref_harness as rh
from helpers import REFERENCE, ROOT, mesh_tri
from oracle import pyoracle as po

TWO_PI = 2.0 * math.pi


def rotating(x, y):
    return (TWO_PI * y, -TWO_PI * x)


need_ref = pytest.mark.skipif(not (rh.available() and os.path.exists(REFERENCE)),
                              reason="reference build oracle/_ref not available")
OWN_TOL = {1: 5e-13, 2: 5e-13, 3: 2e-11, 4: 5e-9}


@need_ref
@pytest.mark.parametrize("periodic", [True, False])
@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_rotating_velocity_residual_matches_the_reference(p, periodic):
    m = rh.square_mesh(5, jitter=0.12, seed=30 + p, periodic=periodic, side_tags=(3, 3, 3, 3))
    o = po.Oracle(2, m["top"], m["xyz"], p, po.LAW_SCALAR, po.FLUX_UPWIND, farfield=(0.3,), bc=m["bc"])
    o.set_velocity_field(rotating)
    U = np.random.default_rng(p).standard_normal(o.ndof)
    Rr, _, _ = rh.run(m, "scalar", "upwind", p, "residual", U, vel="rotate", farfield=(0.3, 0, 0, 0))
    Ro, _ = o.residual(U)
    assert np.abs(Ro - Rr).max() / np.abs(Rr).max() < OWN_TOL[p]


@need_ref
def test_rotating_velocity_rk4_steps_match_the_reference():
    p = 2
    m = rh.square_mesh(6, jitter=0.1, seed=7, periodic=True)
    o = po.Oracle(2, m["top"], m["xyz"], p, po.LAW_SCALAR, po.FLUX_UPWIND, farfield=(0.0,), bc=m["bc"])
    o.set_velocity_field(rotating)
    U = np.random.default_rng(3).standard_normal(o.ndof)
    dt = 0.2 * o.dt(U, 0.1)
    Ur, _, res = rh.run(m, "scalar", "upwind", p, "rk4", U, dt=dt, nsteps=3, vel="rotate", farfield=(0, 0, 0, 0))
    Uo, t = U, 0.0
    for _ in range(3):
        Uo, ro = o.step(Uo, po.RK4, t, dt)
        t += dt
    assert np.abs(Uo - Ur).max() / np.abs(Ur).max() < 1e-12
    assert abs(ro - res[-1]) <= 1e-10 * abs(ro)


def test_constant_field_equals_the_constant_velocity_law():
    vids, xyz, bc = mesh_tri(6, 5, periodic_x=False, periodic_y=True, jitter=0.1, seed=2, side_tags=(4, 4, 4, 4))
    a = po.Oracle(2, vids, xyz, 3, po.LAW_SCALAR, po.FLUX_UPWIND, vel=(1.0, 0.5), farfield=(0.7,), bc=bc)
    b = po.Oracle(2, vids, xyz, 3, po.LAW_SCALAR, po.FLUX_UPWIND, farfield=(0.7,), bc=bc)
    b.set_velocity_field(lambda x, y: (1.0, 0.5))
    U = np.random.default_rng(1).standard_normal(a.ndof)
    assert np.array_equal(a.residual(U)[0], b.residual(U)[0])
    assert a.dt(U, 0.1) == b.dt(U, 0.1)


def test_rotation_conserves_the_total():
    """on a periodic origin-centred square the rotation is divergence free and
    its normal velocity agrees on the two images of every periodic face: the
    residual's element means sum to zero (conservation)"""
    m = rh.square_mesh(8, jitter=0.1, seed=4, periodic=True)
    o = po.Oracle(2, m["top"], m["xyz"], 3, po.LAW_SCALAR, po.FLUX_UPWIND, farfield=(0.0,), bc=m["bc"])
    o.set_velocity_field(rotating)
    U = np.random.default_rng(9).standard_normal(o.ndof)
    R, _ = o.residual(U)
    means = R.reshape(o.E, 10)[:, 0]
    assert abs(means.sum()) <= 1e-12 * np.abs(R).sum()


LIB = os.path.join(ROOT, "tests", "cpp", "libemu.so")


@pytest.mark.skipif(not os.path.exists(LIB), reason="emulator not built")
@pytest.mark.parametrize("form", [0, 2, 1])
@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_emulated_kernel_with_the_rotating_field(p, form):
    """the tiled kernel's phases with the point velocity table (4 and 8 waves,
    FP64; the FP32 form within 5e-5)"""
    lib = ctypes.CDLL(LIB)
    pd = ctypes.POINTER(ctypes.c_double)
    lib.emu_tri_residual_rotating.argtypes = [ctypes.POINTER(dgpar.MeshDesc), ctypes.POINTER(dgpar.LawDesc),
                                              ctypes.c_int, ctypes.c_int, pd, pd]
    vids, xyz, bc = mesh_tri(7, 6, x0=-1, y0=-1, lx=2, ly=2, periodic_x=False, periodic_y=False, jitter=0.1, seed=5,
                             side_tags=(3, 3, 3, 3))
    o = po.Oracle(2, vids, xyz, p, po.LAW_SCALAR, po.FLUX_UPWIND, farfield=(0.4,), bc=bc)
    o.set_velocity_field(rotating)
    U = np.random.default_rng(p).standard_normal(o.ndof)
    md = dgpar.Mesh(2, vids, xyz, bc).desc()
    ld = dgpar.Law(dgpar.LAW_SCALAR, dgpar.FLUX_UPWIND, farfield=(0.4,)).desc()
    R = np.zeros_like(U)
    assert lib.emu_tri_residual_rotating(ctypes.byref(md), ctypes.byref(ld), p, form, U.ctypes.data_as(pd),
                                         R.ctypes.data_as(pd)) == 0
    Ro, _ = o.residual(U)
    tol = 5e-5 if form & 1 else 1e-12
    assert np.abs(R - Ro).max() <= tol * np.abs(Ro).max()


def bump(x, y, z=0.0, t=0.0):
    """ScalarPb's initial condition (CaseOne, Lilia.cc:38-58)"""
    r = math.sqrt((x + 0.5) ** 2 + y * y)
    return (math.cos(TWO_PI * r) ** 2 if r <= 0.25 else 0.0,)


@need_ref
@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_projection_of_a_field_matches_the_reference(p):
    """the oracle's projection of a caller's field against what the reference's
    DGAnalysis constructor projects from its far field (DGAnalysis.cc:200) with
    ScalarPb's bump; the mesh is fine enough that several elements cut the bump"""
    m = rh.square_mesh(8, jitter=0.12, seed=40 + p, periodic=True)
    o = po.Oracle(2, m["top"], m["xyz"], p, po.LAW_SCALAR, po.FLUX_UPWIND, farfield=(0.0,), bc=m["bc"])
    Uo = o.project_fn(bump)
    Ur, _, _ = rh.run(m, "scalar", "upwind", p, "project", np.zeros(o.ndof), vel="rotate", far="bump")
    assert np.abs(Uo).max() > 0.1 and np.count_nonzero(np.abs(Uo) > 1e-3) > 20
    # the reference's printed quadrature tables set the agreement, as for residuals
    assert np.abs(Uo - Ur).max() / np.abs(Ur).max() < OWN_TOL[p]


@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_projection_reproduces_polynomials_of_the_order(p):
    """a field of degree p lies in the element's space: its projection evaluated
    at any point of the element is the field there"""
    vids, xyz, bc = mesh_tri(4, 3, jitter=0.15, seed=2)
    o = po.Oracle(2, vids, xyz, p, po.LAW_EULER2D, po.FLUX_LF, gamma=1.4)
    cs = np.random.default_rng(p).standard_normal((4, p + 1, p + 1))

    def poly(x, y, z=0.0, t=0.0):
        return [sum(cs[f, i, j] * x ** i * y ** j for i in range(p + 1) for j in range(p + 1 - i)) for f in range(4)]
    U = o.project_fn(poly).reshape(o.E, 4, -1)
    rng = np.random.default_rng(9)
    for e in range(0, o.E, 5):
        for _ in range(3):
            a, b = rng.uniform(0, 0.5, 2)
            phi, _ = po.basis(2, p, (a, b, 0))
            X = xyz[e]
            x, y = X[0] + a * (X[1] - X[0]) + b * (X[2] - X[0])
            assert np.allclose(U[e] @ phi, poly(x, y), rtol=0, atol=1e-12)


@need_ref
def test_the_reference_times_its_residual_sweeps():
    """REF_SWEEPS: the driver times its sweeps and still writes one residual"""
    m = rh.square_mesh(4, jitter=0.1, seed=3, periodic=True)
    o = po.Oracle(2, m["top"], m["xyz"], 1, po.LAW_SCALAR, po.FLUX_UPWIND, farfield=(0.0,), bc=m["bc"])
    o.set_velocity_field(rotating)
    U = np.random.default_rng(1).standard_normal(o.ndof)
    stats = {}
    os.environ["REF_SWEEPS"] = "3"
    try:
        Rr, _, _ = rh.run(m, "scalar", "upwind", 1, "residual", U, vel="rotate", stats=stats)
    finally:
        del os.environ["REF_SWEEPS"]
    assert stats["residual_seconds"] > 0
    Ro, _ = o.residual(U)
    assert np.abs(Ro - Rr).max() / np.abs(Rr).max() < OWN_TOL[1]
