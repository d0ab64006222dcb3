"""Maxwell's equations on tetrahedra (eps = mu = 1, perfect-conductor walls).

The reference has Maxwell in 2D only (MaxwellLaw.cc:21-200), so the 3D oracle
is pinned by what the mathematics fixes:
- the Kuhn mesh (6 tetrahedra per box) fills the box, every interior face is
  shared by two tetrahedra and the boundary carries 12 N^2 faces;
- semi-discrete energy: with the central flux and PEC walls the DG operator is
  skew in the mass inner product, sum_e U_e . R_e = 0 to rounding for any state;
  with the upwind flux it is dissipative, sum_e U_e . R_e < 0, and exactly the
  negative of the upwind jump penalty;
- the PEC cavity mode (1,1,1) converges at order p+1 under refinement.
The tiled tet kernel (its phase functions run on the CPU through
tests/cpp/emu_stage.cc) then matches the oracle within 1e-12."""
import ctypes

import numpy as np
import pytest

from oracle import pyoracle as po
from test_kernel_emulation import LIBEMU

PD = ctypes.POINTER(ctypes.c_double)
CAVITY = ()


def kuhn(n):
    import dgpar
    m = dgpar.mesh_tet_structured(n, n, n)
    return m.vids, m.xyz


def oracle(n, p, flux):
    vids, xyz = kuhn(n)
    return vids, xyz, po.Oracle(3, vids, xyz, p, po.LAW_MAXWELL3D, flux)


def test_kuhn_mesh_fills_the_box_and_matches_faces():
    for n in (1, 2, 3):
        vids, xyz = kuhn(n)
        assert vids.shape[0] == 6 * n ** 3
        J = np.stack([xyz[:, 1] - xyz[:, 0], xyz[:, 2] - xyz[:, 0], xyz[:, 3] - xyz[:, 0]], axis=2)
        vol = np.abs(np.linalg.det(J)) / 6
        assert abs(vol.sum() - 1.0) < 1e-13 and vol.min() > 0
        o = po.Oracle(3, vids, xyz, 1, po.LAW_MAXWELL3D, po.FLUX_UPWIND)
        f = o.faces()
        assert (f[:, 2] < 0).sum() == 12 * n * n
        assert 4 * vids.shape[0] == 2 * (f[:, 2] >= 0).sum() + (f[:, 2] < 0).sum()


def mass_weighted_dot(o, U, R):
    return float(np.dot(U, R))  # M = |J| I and dU/dt = R/|J|: d/dt 1/2 U.MU = U.R


@pytest.mark.parametrize("p", [1, 2, 3])
def test_central_flux_is_energy_neutral(p):
    vids, xyz, o = oracle(2, p, po.FLUX_CENTRAL)
    U = np.random.default_rng(p).standard_normal(o.ndof)
    R, _ = o.residual(U)
    assert abs(mass_weighted_dot(o, U, R)) <= 1e-12 * np.abs(U).sum() * np.abs(R).max()


@pytest.mark.parametrize("p", [1, 2, 3])
def test_upwind_flux_dissipates_its_jump_penalty(p):
    vids, xyz, oc = oracle(2, p, po.FLUX_CENTRAL)
    _, _, ou = oracle(2, p, po.FLUX_UPWIND)
    U = np.random.default_rng(10 + p).standard_normal(ou.ndof)
    Ru, _ = ou.residual(U)
    Rc, _ = oc.residual(U)
    du = mass_weighted_dot(ou, U, Ru)
    assert du < 0
    # the upwind terms alone: R_u - R_c carries exactly the dissipation
    assert abs(du - mass_weighted_dot(ou, U, Ru - Rc)) <= 1e-12 * abs(du) * 10


@pytest.mark.parametrize("p", [1, 2])
def test_cavity_mode_converges_at_order_p_plus_one(p):
    errs = []
    for n in (2, 4):
        vids, xyz, o = oracle(n, p, po.FLUX_UPWIND)
        U = o.project(3)
        dt = o.dt(U, 0.1)
        T = 0.05
        steps = int(np.ceil(T / dt))
        dt = T / steps
        t = 0.0
        for _ in range(steps):
            U, _ = o.step(U, po.RK4, t, dt)
            t += dt
        errs.append(np.linalg.norm(o.l2_error(U, 3, (), t)))
    rate = np.log2(errs[0] / errs[1])
    assert rate > p + 1 - 0.6, (errs, rate)


# ---------------------------------------------------------------------------
# the tet kernel through the CPU emulation

if LIBEMU is not None:
    import dgpar

    p_md, p_ld = ctypes.POINTER(dgpar.MeshDesc), ctypes.POINTER(dgpar.LawDesc)
    LIBEMU.emu_tet_residual.argtypes = [p_md, p_ld, ctypes.c_int, ctypes.c_int, PD, PD]
    LIBEMU.emu_tet_steps.argtypes = [p_md, p_ld, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                     PD, PD]
emu = pytest.mark.skipif(LIBEMU is None, reason="emulator not built")


@emu
@pytest.mark.parametrize("shape", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("flux", [po.FLUX_UPWIND, po.FLUX_CENTRAL])
@pytest.mark.parametrize("p", [1, 2, 3])
def test_emulated_tet_residual(flux, p, shape):
    """every tile shape: 4 / 8 waves per 64 tetrahedra with traces over the
    coefficient rows, 8 / 16 waves with the rows kept"""
    vids, xyz, o = oracle(2, p, flux)
    U = np.random.default_rng(20 + p).standard_normal(o.ndof)
    md = dgpar.Mesh(3, vids, xyz).desc()
    ld = dgpar.MaxwellLaw(flux).desc()
    R = np.zeros_like(U)
    assert LIBEMU.emu_tet_residual(ctypes.byref(md), ctypes.byref(ld), p, shape, U.ctypes.data_as(PD),
                                   R.ctypes.data_as(PD)) == 0
    Ro, _ = o.residual(U)
    assert np.abs(R - Ro).max() <= 1e-12 * np.abs(Ro).max()


@emu
@pytest.mark.parametrize("shape", [0, 1, 2, 3, 4])
def test_emulated_tet_rk4_steps(shape):
    p = 3
    vids, xyz, o = oracle(2, p, po.FLUX_UPWIND)
    U = o.project(3)
    dt = o.dt(U, 0.1)
    Uo, t = U, 0.0
    for _ in range(3):
        Uo, ro = o.step(Uo, po.RK4, t, dt)
        t += dt
    md = dgpar.Mesh(3, vids, xyz).desc()
    ld = dgpar.MaxwellLaw(po.FLUX_UPWIND).desc()
    Ue = U.copy()
    r = ctypes.c_double()
    assert LIBEMU.emu_tet_steps(ctypes.byref(md), ctypes.byref(ld), p, shape, po.RK4, dt, 3, Ue.ctypes.data_as(PD),
                                ctypes.byref(r)) == 0
    assert np.abs(Ue - Uo).max() <= 1e-12 * np.abs(Uo).max()
    assert abs(r.value - ro) <= 1e-10 * abs(ro)


if LIBEMU is not None:
    LIBEMU.emu_tet_steps32.argtypes = [p_md, p_ld, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                       ctypes.c_int, PD]


@emu
@pytest.mark.parametrize("shape", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("p", [1, 2, 3])
def test_emulated_tet_fp32_form_rk4_steps(p, shape):
    """the FP32 form (FP32 arithmetic, tables and stage copies; FP64 step-start
    state, accumulator and update): 3 RK4 steps of the cavity mode against the
    FP64 oracle within FP32 rounding (2e-6 of the field scale), and closer to
    the oracle than the step itself moves the state"""
    vids, xyz, o = oracle(2, p, po.FLUX_UPWIND)
    U = o.project(3)
    dt = o.dt(U, 0.1)
    Uo, t = U, 0.0
    for _ in range(3):
        Uo, _ = o.step(Uo, po.RK4, t, dt)
        t += dt
    md = dgpar.Mesh(3, vids, xyz).desc()
    ld = dgpar.MaxwellLaw(po.FLUX_UPWIND).desc()
    Ue = U.copy()
    assert LIBEMU.emu_tet_steps32(ctypes.byref(md), ctypes.byref(ld), p, shape, po.RK4, dt, 3,
                                  Ue.ctypes.data_as(PD)) == 0
    err = np.abs(Ue - Uo).max() / np.abs(Uo).max()
    assert err <= 2e-6
    assert err < 1e-2 * np.abs(Uo - U).max() / np.abs(Uo).max()


@emu
@pytest.mark.parametrize("shape", [0, 2, 4])
def test_emulated_tet_fp32_first_stage_from_rounded_state(shape):
    """The library's FP32 step sequence on tetrahedra: the last stage also writes
    the new state rounded to FP32, and the next step's first stage reads that copy
    through the FP32-input kernel.  The FP64-input kernel rounds the same values
    on load, so the steps are bit for bit those of the FP64-input first stage."""
    vids, xyz, o = oracle(2, 2, po.FLUX_UPWIND)
    U = o.project(3)
    dt = o.dt(U, 0.1)
    md = dgpar.Mesh(3, vids, xyz).desc()
    ld = dgpar.MaxwellLaw(po.FLUX_UPWIND).desc()
    out = []
    for m32 in (0, 1):
        LIBEMU.emu_tet_m32(m32)
        Ue = U.copy()
        try:
            assert LIBEMU.emu_tet_steps32(ctypes.byref(md), ctypes.byref(ld), 2, shape, po.RK4, dt, 3,
                                          Ue.ctypes.data_as(PD)) == 0
        finally:
            LIBEMU.emu_tet_m32(0)
        out.append(Ue)
    np.testing.assert_array_equal(out[0], out[1])
    assert np.abs(out[1] - U).max() > 0
