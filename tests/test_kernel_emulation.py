"""CPU emulation of the tiled HIP stage kernel (its own phase functions run
thread by thread on the host, tests/cpp/emu_stage.cc) against the oracle:
FP64 residuals and RK steps within 1e-12; the FP32 form within its budget.
Catches arithmetic and indexing errors in the kernel code before it reaches a
GPU."""
import ctypes
import os

import numpy as np
import pytest

import dgpar
from helpers import ROOT, VORTEX, mesh_tri
from oracle import pyoracle as po

LIB = os.path.join(ROOT, "tests", "cpp", "libemu.so")
pytestmark = pytest.mark.skipif(not os.path.exists(LIB), reason="emulator not built")


def _lib():
    lib = ctypes.CDLL(LIB)
    p_md, p_ld = ctypes.POINTER(dgpar.MeshDesc), ctypes.POINTER(dgpar.LawDesc)
    pd = ctypes.POINTER(ctypes.c_double)
    lib.emu_tri_residual.argtypes = [p_md, p_ld, ctypes.c_int, ctypes.c_int, pd, pd]
    lib.emu_tri_steps.argtypes = [p_md, p_ld, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                  ctypes.c_int, pd, pd]
    return lib


LIBEMU = _lib() if os.path.exists(LIB) else None
FF = (1.0, 1.0, 1.0, 3.5)


def case(law, flux, p, periodic=True, nx=8, ff=FF, tags=(40000, 7, 40000, 7)):
    if law == "scalar":
        vids, xyz, bc = mesh_tri(nx, nx, periodic_x=periodic, periodic_y=periodic, jitter=0.1, seed=3, side_tags=tags)
        o = po.Oracle(2, vids, xyz, p, po.LAW_SCALAR, po.FLUX_UPWIND, vel=(1.0, 0.5), farfield=(0.2,), bc=bc)
        lw = dgpar.Law(dgpar.LAW_SCALAR, dgpar.FLUX_UPWIND, vel=(1.0, 0.5), farfield=(0.2,))
        U = o.project(0)
    else:
        vids, xyz, bc = mesh_tri(nx, nx, lx=10, ly=10, periodic_x=periodic, periodic_y=periodic, jitter=0.1, seed=3,
                                 side_tags=tags)
        o = po.Oracle(2, vids, xyz, p, po.LAW_EULER2D, po.FLUX_LF if flux == "lf" else po.FLUX_ROE, gamma=1.4,
                      farfield=ff, bc=bc)
        lw = dgpar.Law(dgpar.LAW_EULER2D, dgpar.FLUX_LF if flux == "lf" else dgpar.FLUX_ROE, gamma=1.4, farfield=ff)
        U = o.project(1, VORTEX)
    return dgpar.Mesh(2, vids, xyz, bc), lw, o, U


def emu_residual(mesh, lw, p, U, fp32=0):
    md, ld = mesh.desc(), lw.desc()
    R = np.zeros_like(U)
    U = np.ascontiguousarray(U)
    pd = ctypes.POINTER(ctypes.c_double)
    assert LIBEMU.emu_tri_residual(ctypes.byref(md), ctypes.byref(ld), p, fp32, U.ctypes.data_as(pd),
                                   R.ctypes.data_as(pd)) == 0
    return R


def emu_steps(mesh, lw, p, U, scheme, dt, n, fp32=0):
    md, ld = mesh.desc(), lw.desc()
    U = np.array(U, copy=True)
    r = ctypes.c_double()
    pd = ctypes.POINTER(ctypes.c_double)
    assert LIBEMU.emu_tri_steps(ctypes.byref(md), ctypes.byref(ld), p, fp32, scheme, dt, n, U.ctypes.data_as(pd),
                                ctypes.byref(r)) == 0
    return U, r.value


def rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.mark.parametrize("law,flux,periodic", [("scalar", None, True), ("scalar", None, False), ("euler", "lf", True),
                                               ("euler", "roe", True), ("euler", "lf", False), ("euler", "roe", False)])
@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_emulated_residual(law, flux, periodic, p):
    mesh, lw, o, U = case(law, flux, p, periodic)
    assert rel(emu_residual(mesh, lw, p, U), o.residual(U)[0]) < 1e-12


@pytest.mark.parametrize("ff", [FF, (1.0, 2.0, 0.5, 4.625)])
@pytest.mark.parametrize("flux", ["lf", "roe"])
@pytest.mark.parametrize("p", [1, 3])
def test_emulated_characteristic_farfield(ff, flux, p):
    """tag 30000 (EulerLaw.cc:185-278) on three sides, a wall on the fourth:
    the kernel's boundary state under the law's own flux against the oracle's
    direct flux formulas; residual, then RK4 steps"""
    mesh, lw, o, U = case("euler", flux, p, periodic=False, ff=ff, tags=(30000, 30000, 40000, 30000))
    Ro = o.residual(U)[0]
    assert rel(emu_residual(mesh, lw, p, U), Ro) < 1e-12
    assert rel(emu_residual(mesh, lw, p, U, fp32=1), Ro) < 5e-5
    dt = o.dt(U, 0.1)
    Uo, t = U, 0.0
    for s in range(2):
        Uo, _ = o.step(Uo, po.RK4, t, dt)
        t += dt
    assert rel(emu_steps(mesh, lw, p, U, po.RK4, dt, 2)[0], Uo) < 1e-12


@pytest.mark.parametrize("law,flux,p,scheme", [("scalar", None, 2, po.RK_TVD3), ("scalar", None, 1, po.RK_FE),
                                               ("euler", "lf", 3, po.RK4), ("euler", "roe", 2, po.RK_TVD2)])
def test_emulated_steps(law, flux, p, scheme):
    mesh, lw, o, U = case(law, flux, p, periodic=(law == "scalar"))
    dt = o.dt(U, 0.1)
    Uo, t = U, 0.0
    for s in range(3):
        Uo, ro = o.step(Uo, scheme, t, dt)
        t += dt
    Ue, re = emu_steps(mesh, lw, p, U, scheme, dt, 3)
    assert rel(Ue, Uo) < 1e-12
    assert abs(re - ro) <= 1e-10 * abs(ro)


@pytest.mark.parametrize("law,flux,periodic", [("scalar", None, True), ("euler", "lf", True), ("euler", "roe", False)])
@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_emulated_residual_8_wave_tiles(law, flux, periodic, p):
    """the 8-wave tile shape (512 threads per 64 elements) computes the same rows"""
    mesh, lw, o, U = case(law, flux, p, periodic)
    assert rel(emu_residual(mesh, lw, p, U, fp32=2), o.residual(U)[0]) < 1e-12


def test_emulated_8_wave_steps_and_fp32_form():
    mesh, lw, o, U = case("euler", "roe", 3, periodic=False)
    dt = o.dt(U, 0.1)
    Uo, t = U, 0.0
    for s in range(3):
        Uo, ro = o.step(Uo, po.RK4, t, dt)
        t += dt
    Ue, re = emu_steps(mesh, lw, 3, U, po.RK4, dt, 3, fp32=2)
    assert rel(Ue, Uo) < 1e-12
    assert abs(re - ro) <= 1e-10 * abs(ro)
    U32, _ = emu_steps(mesh, lw, 3, U, po.RK4, dt, 3, fp32=3)
    assert rel(U32, Uo) < 1e-5


@pytest.mark.parametrize("law,flux,p,scheme,shape", [("euler", "lf", 3, po.RK4, 0), ("euler", "roe", 2, po.RK_TVD3, 2),
                                                     ("scalar", None, 1, po.RK_TVD2, 0), ("euler", "lf", 4, po.RK4, 2)])
def test_fp32_form_first_stage_from_rounded_state(law, flux, p, scheme, shape):
    """The library's FP32 step sequence: the last stage also writes the new state
    rounded to FP32, and the next step's first stage reads that copy through the
    FP32-input kernel.  The FP64-input kernel rounds the same values on load, so
    the steps are bit for bit those of the FP64-input first stage."""
    mesh, lw, o, U = case(law, flux, p, periodic=(law == "scalar"))
    dt = o.dt(U, 0.1)
    Ua, ra = emu_steps(mesh, lw, p, U, scheme, dt, 3, fp32=1 | shape)
    Ub, rb = emu_steps(mesh, lw, p, U, scheme, dt, 3, fp32=1 | shape | 4)
    np.testing.assert_array_equal(Ua, Ub)
    assert ra == rb
    # the FP32 form is not the FP64 one (the comparison above is not vacuous)
    U64, _ = emu_steps(mesh, lw, p, U, scheme, dt, 3)
    assert 0 < rel(Ub, U64) < 1e-5


def test_emulated_fp32_form():
    mesh, lw, o, U = case("euler", "lf", 3, nx=12)
    R64 = o.residual(U)[0]
    assert rel(emu_residual(mesh, lw, 3, U, fp32=1), R64) < 1e-5
    dt = o.dt(U, 0.1)
    U64, _ = emu_steps(mesh, lw, 3, U, po.RK4, dt, 5)
    U32, _ = emu_steps(mesh, lw, 3, U, po.RK4, dt, 5, fp32=1)
    assert rel(U32, U64) < 1e-5
