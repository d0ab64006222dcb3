"""Maxwell2d and Burgers2D (MaxwellLaw.cc:21-200, Burgers2D.cc:22-92) through the
kernels' code run on the CPU (tests/cpp/emu_stage.cc), against the oracle, whose
restatement of both laws is pinned against the reference build in
test_oracle_vs_reference.py.

- the tiled stage kernel (4- and 8-wave tiles) and its FP32 form: residuals of
  both laws on periodic meshes and on meshes with walls and far-field sides;
- RK steps;
- the mixed-order persistent kernel (P1-P4 in one launch), whose layouts depend
  on the field count (3 fields is new);
- the derivative kernels: JVP against finite differences of the oracle, and the
  dot-product identity of JVP and VJP;
- the boundary fluxes: Maxwell2d walls mirror H about the normal (the energy
  flux through a wall is zero), Burgers2D takes the far field on inflow only."""
import ctypes

import numpy as np
import pytest

from helpers import mesh_tri, np_tri
from oracle import pyoracle as po
from test_kernel_emulation import LIBEMU, emu_residual, emu_steps, pytestmark  # noqa: F401

PD = ctypes.POINTER(ctypes.c_double)
FF = {"maxwell2d": (0.3, -0.2, 0.5), "burgers": (0.4,)}


def case(law, p, periodic=True, nx=8, orders=None):
    import dgpar
    vids, xyz, bc = mesh_tri(nx, nx, periodic_x=periodic, periodic_y=periodic, jitter=0.1, seed=5,
                             side_tags=(40000, 7, 40000, 7))
    lw_o = po.LAW_MAXWELL2D if law == "maxwell2d" else po.LAW_BURGERS2D
    fl = po.FLUX_UPWIND if law == "maxwell2d" else po.FLUX_CENTRAL
    o = po.Oracle(2, vids, xyz, p if orders is None else orders, lw_o, fl, farfield=FF[law], bc=bc)
    U = 0.5 * np.random.default_rng(p).standard_normal(o.ndof)
    lw = dgpar.Maxwell2dLaw(FF[law]) if law == "maxwell2d" else dgpar.Burgers2DLaw(FF[law])
    return dgpar.Mesh(2, vids, xyz, bc, order=orders), lw, o, U


def rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.mark.parametrize("law", ["maxwell2d", "burgers"])
@pytest.mark.parametrize("periodic", [True, False])
@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_emulated_residual(law, periodic, p):
    mesh, lw, o, U = case(law, p, periodic)
    Ro = o.residual(U)[0]
    assert rel(emu_residual(mesh, lw, p, U), Ro) < 1e-12
    assert rel(emu_residual(mesh, lw, p, U, fp32=2), Ro) < 1e-12  # 8-wave tiles
    assert rel(emu_residual(mesh, lw, p, U, fp32=1), Ro) < 5e-5   # FP32 form


@pytest.mark.parametrize("law", ["maxwell2d", "burgers"])
@pytest.mark.parametrize("periodic", [True, False])
@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_emulated_element_kernel_residual(law, periodic, p):
    """the element-per-work-item kernel's code (dg_element.h), run on the CPU:
    the library refuses this kernel for these laws on the GPU (Maxwell2d gave
    wrong values there at P3 and P4), while the same source on the host matches
    the oracle"""
    mesh, lw, o, U = case(law, p, periodic)
    md, ld = mesh.desc(), lw.desc()
    R = np.zeros_like(U)
    assert LIBEMU.emu_tri_residual_element(ctypes.byref(md), ctypes.byref(ld), p, U.ctypes.data_as(PD),
                                           R.ctypes.data_as(PD)) == 0
    assert rel(R, o.residual(U)[0]) < 1e-12


@pytest.mark.parametrize("law,p,scheme", [("maxwell2d", 3, po.RK4), ("burgers", 2, po.RK_TVD3)])
def test_emulated_steps(law, p, scheme):
    mesh, lw, o, U = case(law, p, periodic=False)
    dt = o.dt(U, 0.1)
    Uo, t = U, 0.0
    for s in range(3):
        Uo, ro = o.step(Uo, scheme, t, dt)
        t += dt
    Ue, re = emu_steps(mesh, lw, p, U, scheme, dt, 3)
    assert rel(Ue, Uo) < 1e-12
    assert abs(re - ro) <= 1e-10 * abs(ro)


@pytest.mark.parametrize("law", ["maxwell2d", "burgers"])
def test_emulated_mixed_orders(law):
    import dgpar
    orders = np.random.default_rng(9).integers(1, 5, 2 * 10 * 10).astype(np.int32)
    mesh, lw, o, U = case(law, 0, periodic=False, nx=10, orders=orders)
    md, ld = mesh.desc(), lw.desc()
    LIBEMU.emu_mixed_residual.argtypes = [ctypes.POINTER(dgpar.MeshDesc), ctypes.POINTER(dgpar.LawDesc), PD, PD]
    R = np.zeros_like(U)
    assert LIBEMU.emu_mixed_residual(ctypes.byref(md), ctypes.byref(ld), U.ctypes.data_as(PD), R.ctypes.data_as(PD)) == 0
    assert rel(R, o.residual(U)[0]) < 1e-12


@pytest.mark.parametrize("law", ["maxwell2d", "burgers"])
@pytest.mark.parametrize("p", [1, 3])
def test_emulated_derivatives(law, p):
    from test_adjoint_emulation import emu_derivative
    mesh, lw, o, U = case(law, p, periodic=False)
    V = np.random.default_rng(40 + p).standard_normal(U.shape)
    h = 1e-6
    Av = emu_derivative(mesh, lw, p, U, V, True)
    fd = (o.residual(U + h * V)[0] - o.residual(U - h * V)[0]) / (2 * h)
    assert np.abs(Av - fd).max() <= 1e-6 * np.abs(fd).max()
    W = np.random.default_rng(50 + p).standard_normal(U.shape)
    ATw = emu_derivative(mesh, lw, p, U, W, False, nd=1)
    assert abs(float(np.dot(Av, W)) - float(np.dot(V, ATw))) <= 1e-12 * np.abs(Av).sum() * np.abs(W).max()


def test_maxwell2d_walls_conserve_energy_flux():
    """With walls on every side, sum_e U^T M R(U) (M = |J| I) is the energy flux
    through the boundary plus the interior upwind dissipation (<= 0): the wall
    mirror adds no energy, and without the interior jumps (a continuous linear
    field) the balance is zero to rounding."""
    vids, xyz, bc = mesh_tri(6, 6, periodic_x=False, periodic_y=False, jitter=0.0, side_tags=(40000,) * 4)
    o = po.Oracle(2, vids, xyz, 1, po.LAW_MAXWELL2D, po.FLUX_UPWIND, bc=bc)
    U = 0.5 * np.random.default_rng(3).standard_normal(o.ndof)
    R = o.residual(U)[0]
    dj = np.repeat(o.detj(), 3 * np_tri(1))
    assert float(np.dot(U, dj * R)) <= 1e-12
    # a field with tangential H = 0 and Ez = 0 on every wall of the unit square:
    # H = (0, 0) everywhere, Ez = 0 -> zero residual; Ez = const -> zero flux
    Uc = np.zeros(o.ndof)
    Uc.reshape(-1, 3, np_tri(1))[:, 2, 0] = 0.7 / np.sqrt(2.0)
    assert np.abs(o.residual(Uc)[0]).max() <= 1e-13


def test_burgers_boundary_upwinds_the_far_field():
    """inflow faces (nx + ny < 0: the left and bottom sides) take the far-field
    flux and outflow faces the interior flux: a constant state equal to the far
    field has zero residual, another constant state only changes on inflow
    elements"""
    vids, xyz, bc = mesh_tri(4, 4, periodic_x=False, periodic_y=False, jitter=0.0, side_tags=(7,) * 4)
    o = po.Oracle(2, vids, xyz, 1, po.LAW_BURGERS2D, po.FLUX_CENTRAL, farfield=(0.4,), bc=bc)
    U = np.zeros(o.ndof)
    U[::3] = 0.4 / np.sqrt(2.0)  # the constant 0.4: phi_0 = sqrt(2)
    assert np.abs(o.residual(U)[0]).max() <= 1e-15
    U[::3] = 0.9 / np.sqrt(2.0)
    R = o.residual(U)[0].reshape(-1, 3)
    touched = np.abs(R).max(axis=1) > 1e-12
    cx = xyz[:, :, 0].mean(axis=1)
    cy = xyz[:, :, 1].mean(axis=1)
    inflow_edge = (cx < 0.25) | (cy < 0.25)
    assert touched.any() and not (touched & ~inflow_edge).any()
