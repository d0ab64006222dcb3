"""Local time stepping: the oracle's restatement (oracle/dg_oracle.c or_lts_*)
pinned by what the method's definition fixes, and the kernels' stage skipping
(CPU emulation of the tiled stage kernel over the library's class schedule)
against the oracle, which evaluates every stage of every element.

The reference's RungeKuttaTVD2adaptive (TimeIntegrator.cc:442-833) calls range
assembly overloads that are empty (TimeIntegrator.cc:1438-1487), so there is no
reference output to compare with; the level rule follows its size bins
(DGAnalysis.cc:749-890).  What pins the oracle:
  * one level: the step equals SSP-RK2 (RungeKuttaTVD2, TimeIntegrator.cc:279-438);
  * a partitioned Runge-Kutta step built here from Butcher tableaux derived
    independently (stage states relative to the step start) gives the same state;
  * the tableaux satisfy the second-order conditions, and the measured order of
    the time error on a graded mesh is 2;
  * totals are conserved to rounding and a free stream stays constant.
"""
import ctypes
import os

import numpy as np
import pytest

from helpers import ROOT, graded_nodes, mesh_tri, mesh_tri_graded
from oracle import pyoracle as po

GRADE = (4, 2, 1, 1, 2, 4)


def graded(n=12, lo=0.0, hi=1.0):
    xs = graded_nodes(n, lo, hi, GRADE)
    return mesh_tri_graded(xs, xs)


def totals(o, U):
    """sum_e |J_e| U[e][f][0]: proportional to the integral of each field"""
    detj = o.detj()
    Np = o.ndof // (o.E * o.nfields)
    return (U.reshape(o.E, o.nfields, Np)[:, :, 0] * detj[:, None]).sum(axis=0)


def scalar_oracle(vids, xyz, p):
    return po.Oracle(2, vids, xyz, p, po.LAW_SCALAR, po.FLUX_UPWIND, vel=(1.0, 0.5))


def euler_oracle(vids, xyz, p, flux=po.FLUX_LF):
    return po.Oracle(2, vids, xyz, p, po.LAW_EULER2D, flux, gamma=1.4)


def bins_direct(size, max_levels):
    """the reference's bins (DGAnalysis.cc:772-803), written as a closed form"""
    smax = size.max()
    b = np.floor(np.log2(1.0001 * smax / size) + 1e-12).astype(int)
    return np.minimum(b, max_levels - 1)


def tri_size(xyz):
    a = np.linalg.norm(xyz[:, 1] - xyz[:, 0], axis=1)
    b = np.linalg.norm(xyz[:, 2] - xyz[:, 1], axis=1)
    c = np.linalg.norm(xyz[:, 0] - xyz[:, 2], axis=1)
    s = 0.5 * (a + b + c)
    return 2.0 * np.sqrt(s * (s - a) * (s - b) * (s - c)) / s


def test_levels_follow_the_size_bins():
    vids, xyz = graded()
    o = scalar_oracle(vids, xyz, 1)
    for ml in (1, 2, 3, 4):
        lev, n = o.lts_levels(ml)
        ref = bins_direct(tri_size(xyz), ml)
        F = o.faces()
        inner = F[F[:, 2] >= 0]
        if np.abs(ref[inner[:, 0]] - ref[inner[:, 2]]).max() <= 1:
            np.testing.assert_array_equal(lev, ref)
        assert n == lev.max() + 1 and lev.max() <= ml - 1
        assert np.abs(lev[inner[:, 0]] - lev[inner[:, 2]]).max() <= 1
    assert o.lts_levels(4)[1] >= 3


def test_levels_are_smoothed_to_one_level_across_faces():
    # a jump of 8x in cell width: bins two levels apart meet at a face
    xs = np.concatenate([np.linspace(0, 0.5, 5), np.linspace(0.5, 1.0, 33)[1:]])
    ys = np.linspace(0, 1, 9)
    vids, xyz = mesh_tri_graded(xs, ys)
    o = scalar_oracle(vids, xyz, 1)
    lev, n = o.lts_levels(6)
    ref = bins_direct(tri_size(xyz), 6)
    F = o.faces()
    inner = F[F[:, 2] >= 0]
    assert np.abs(ref[inner[:, 0]] - ref[inner[:, 2]]).max() >= 2
    assert np.abs(lev[inner[:, 0]] - lev[inner[:, 2]]).max() <= 1
    assert (lev >= ref).all()  # only ever moved to a finer level


@pytest.mark.parametrize("p", [1, 2])
def test_one_level_is_ssp_rk2(p):
    vids, xyz, _ = mesh_tri(6, 6, jitter=0.1, seed=5)
    o = scalar_oracle(vids, xyz, p)
    U = o.project(0)
    lev, n = o.lts_levels(1)
    assert n == 1 and not lev.any()
    h = o.dt(U, 0.1)
    Ul = o.lts_step(U, lev, h)
    Ur, _ = o.step(U, po.RK_TVD2, 0.0, h)
    assert np.abs(Ul - Ur).max() <= 1e-14 * np.abs(Ur).max()


def prk_tableau(nlev, b):
    """Butcher matrix (S x S) and weights of level b, from the substep picture:
    M = 2^b Heun substeps of h/M, each spanning L = S/M stages; relative to the
    step start y_n, the state at the start of substep m is y_n + (h/S) sum of
    every earlier stage's k, and an odd stage adds (h/M) k of the stage before."""
    S = 1 << nlev
    M = 1 << b
    L = S // M
    A = np.zeros((S, S))
    for i in range(S):
        m, j = divmod(i, L)
        A[i, : m * L] = 1.0 / S
        if j % 2 == 1:
            A[i, i - 1] += 1.0 / M
    return A, np.full(S, 1.0 / S)


def test_tableaux_are_second_order():
    for nlev in (1, 2, 3, 4):
        S = 1 << nlev
        for b in range(nlev):
            A, w = prk_tableau(nlev, b)
            c = A.sum(axis=1)
            assert abs(w.sum() - 1.0) < 1e-15
            assert abs(w @ c - 0.5) < 1e-15
            # explicit: strictly lower triangular
            assert not np.triu(A).any()
        if nlev == 1:
            A, w = prk_tableau(1, 0)
            np.testing.assert_allclose(A, [[0, 0], [1, 0]])  # Heun
            np.testing.assert_allclose(w, [0.5, 0.5])


def prk_step(o, lev, U, h):
    """one step of the partitioned Runge-Kutta method from its tableaux, with the
    oracle's residual as the only shared piece"""
    nlev = lev.max() + 1
    S = 1 << nlev
    tabs = [prk_tableau(nlev, b) for b in range(nlev)]
    Np = o.ndof // (o.E * o.nfields)
    detj = o.detj()
    Ue = U.reshape(o.E, -1)
    K = []
    for i in range(S):
        Y = Ue.copy()
        for j in range(i):
            a = np.array([tabs[lev[e]][0][i, j] for e in range(o.E)])
            Y += h * a[:, None] * K[j]
        R, _ = o.residual(Y.ravel())
        K.append(R.reshape(o.E, -1) / detj[:, None])
    out = Ue.copy()
    for i in range(S):
        out += h * (1.0 / S) * K[i]
    return out.ravel()


@pytest.mark.parametrize("law", ["scalar", "euler"])
def test_step_equals_the_partitioned_rk_tableau(law):
    vids, xyz = graded(12, 0.0, 10.0) if law == "euler" else graded(12)
    if law == "euler":
        o = euler_oracle(vids, xyz, 2)
        U = o.project(1, (5.0, 5.0, 5.0, 1.0, 1.0, 10.0))
    else:
        o = scalar_oracle(vids, xyz, 2)
        U = o.project(0)
    lev, n = o.lts_levels(3)
    assert n == 3
    h = o.lts_dt(U, 0.1, lev)
    Ul = o.lts_step(U, lev, h)
    Up = prk_step(o, lev, U, h)
    assert np.abs(Ul - Up).max() <= 1e-13 * np.abs(Up).max()
    assert np.abs(Ul - U).max() > 1e-4 * np.abs(U).max()  # it moved


def test_conservation_and_free_stream():
    vids, xyz = graded(12, 0.0, 10.0)
    o = euler_oracle(vids, xyz, 2, po.FLUX_ROE)
    U = o.project(1, (5.0, 5.0, 5.0, 1.0, 1.0, 10.0))
    lev, n = o.lts_levels(3)
    h = o.lts_dt(U, 0.1, lev)
    t0 = totals(o, U)
    V = U
    for _ in range(5):
        V = o.lts_step(V, lev, h)
    np.testing.assert_allclose(totals(o, V), t0, rtol=1e-13, atol=1e-13)
    # a free stream stays constant
    W = o.project(1, (0.0, 5.0, 5.0, 1.0, 1.0, 10.0))  # vortex strength 0
    W2 = o.lts_step(W, lev, h)
    assert np.abs(W2 - W).max() <= 1e-13 * np.abs(W).max()


def test_time_error_is_second_order():
    vids, xyz = graded(12)
    o = scalar_oracle(vids, xyz, 2)
    U = o.project(0)
    lev, n = o.lts_levels(3)
    h0 = o.lts_dt(U, 0.2, lev)
    T = 8 * h0
    # reference: the same semi-discretisation with RK4 at a small global step
    dt = h0 / 64.0
    Ur = U.copy()
    for k in range(8 * 64):
        Ur, _ = o.step(Ur, po.RK4, k * dt, dt)
    errs = []
    for m in (1, 2, 4):
        h = h0 / m
        V = U.copy()
        for _ in range(8 * m):
            V = o.lts_step(V, lev, h)
        errs.append(np.abs(V - Ur).max())
    r1, r2 = errs[0] / errs[1], errs[1] / errs[2]
    assert 3.2 < r1 < 5.0 and 3.5 < r2 < 4.6, errs


# ---- the kernels' stage skipping, CPU emulation against the oracle ------------
LIBEMU_PATH = os.path.join(ROOT, "tests", "cpp", "libemu.so")


def emu_lts(vids, xyz, law, p, max_levels, h, nsteps, U):
    """nsteps steps through the library's planning and the stage kernel's phase
    functions run on the host (tests/cpp/emu_stage.cc emu_tri_lts_steps)"""
    if not os.path.exists(LIBEMU_PATH):
        pytest.skip("tests/cpp/libemu.so not built")
    import dgpar
    lib = ctypes.CDLL(LIBEMU_PATH)
    pd, pi = ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int32)
    lib.emu_tri_lts_steps.argtypes = [ctypes.POINTER(dgpar.MeshDesc), ctypes.POINTER(dgpar.LawDesc), ctypes.c_int,
                                      ctypes.c_int, ctypes.c_double, ctypes.c_int, pd, pi]
    m = dgpar.Mesh(2, vids, xyz)
    md, ld = m.desc(), law.desc()
    V = np.array(U, np.float64, copy=True)
    lev = np.zeros(len(vids), np.int32)
    assert lib.emu_tri_lts_steps(ctypes.byref(md), ctypes.byref(ld), p, max_levels, h, nsteps,
                                 V.ctypes.data_as(pd), lev.ctypes.data_as(pi)) == 0
    return V, lev


@pytest.mark.parametrize("law,p,ml", [("scalar", 2, 3), ("scalar", 1, 4), ("euler", 1, 3), ("euler", 3, 2)])
def test_emulated_kernel_schedule_matches_oracle(law, p, ml):
    import dgpar
    if law == "euler":
        vids, xyz = graded(12, 0.0, 10.0)
        o = euler_oracle(vids, xyz, p)
        U = o.project(1, (5.0, 5.0, 5.0, 1.0, 1.0, 10.0))
        lw = dgpar.EulerLaw(dgpar.FLUX_LF, 1.4)
    else:
        vids, xyz = graded(12)
        o = scalar_oracle(vids, xyz, p)
        U = o.project(0)
        lw = dgpar.Law(dgpar.LAW_SCALAR, dgpar.FLUX_UPWIND, vel=(1.0, 0.5), farfield=(0.0,))
    lev, n = o.lts_levels(ml)
    h = o.lts_dt(U, 0.1, lev)
    V = U
    for _ in range(2):
        V = o.lts_step(V, lev, h)
    W, elev = emu_lts(vids, xyz, lw, p, ml, h, 2, U)
    np.testing.assert_array_equal(elev, lev)
    assert np.abs(W - V).max() <= 1e-12 * np.abs(V).max()
