"""The host C++ mirror classes (ConservationLaw / DGAnalysis / TimeIntegrator
over the C ABI) driven from a C++ program, checked against the CPU oracle."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from helpers import ROOT, VORTEX, mesh_tri
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu
EXE = os.path.join(ROOT, "tests", "cpp", "dropin_vortex")


@pytest.mark.skipif(not os.path.exists(EXE), reason="C++ driver not built")
def test_cpp_dropin_matches_oracle():
    print("\n[gpu case] C++ drop-in vortex", flush=True)
    with tempfile.TemporaryDirectory() as d:
        pre = os.path.join(d, "run")
        r = subprocess.run([EXE, pre], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        U0 = np.fromfile(pre + ".U0")
        R = np.fromfile(pre + ".R")
        U2 = np.fromfile(pre + ".U2")
        lines = dict(l.split(None, 1) for l in r.stdout.splitlines())
        dt = float(lines["dt"])
    vids, xyz, bc = mesh_tri(10, 10, lx=10.0, ly=10.0, jitter=0.1, seed=77)
    # dgpar::Euler2d carries the reference's own flux, the exact Godunov flux
    o = po.Oracle(2, vids, xyz, 3, po.LAW_EULER2D, po.FLUX_EXACT, gamma=1.4, farfield=(1.0, 1.0, 1.0, 3.5))
    Uo = o.project(1, VORTEX)
    assert np.abs(U0 - Uo).max() / np.abs(Uo).max() < 1e-13
    Ro, _ = o.residual(U0)
    assert np.abs(R - Ro).max() / np.abs(Ro).max() < 1e-12
    assert abs(dt - o.dt(U0, 0.1)) <= 1e-14 * dt
    U = U0
    for s in range(2):
        U, _ = o.step(U, po.RK4, s * dt, dt)
    assert np.abs(U2 - U).max() / np.abs(U).max() < 1e-12


def run_exe(mode):
    d = tempfile.mkdtemp()
    pre = os.path.join(d, "run")
    r = subprocess.run([EXE, pre, mode], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = dict(l.split(None, 1) for l in r.stdout.splitlines())
    return pre, float(lines["dt"]), [float(x) for x in lines["resid"].split()]


@pytest.mark.skipif(not os.path.exists(EXE), reason="C++ driver not built")
def test_cpp_dropin_shock_tube_with_limiter_and_adaptation():
    print("\n[gpu case] C++ drop-in shock tube", flush=True)
    pre, dt, res = run_exe("tube")
    U0, R, U2, U3 = (np.fromfile(pre + x) for x in (".U0", ".R", ".U2", ".U3"))
    O0 = np.fromfile(pre + ".O0").astype(np.int32)
    O = np.fromfile(pre + ".O").astype(np.int32)
    vids, xyz, bc = mesh_tri(16, 4, lx=1.0, ly=0.25, periodic_x=False, periodic_y=False, jitter=0.1, seed=5,
                             side_tags=(40000, 2, 40000, 1))
    ff = (1.0, 0.0, 0.0, 2.5, 0.125, 0.0, 0.0, 0.25)
    o = po.Oracle(2, vids, xyz, O0, po.LAW_EULER2D, po.FLUX_ROE_EF, gamma=1.4, farfield=ff, bc=bc)
    Uo, _ = o.limit(o.project(2, (0.53,)))
    assert np.abs(U0 - Uo).max() / np.abs(Uo).max() < 1e-13
    Ro, _ = o.residual(U0)
    assert np.abs(R - Ro).max() / np.abs(Ro).max() < 1e-12
    assert abs(dt - o.dt(U0, 0.1)) <= 1e-13 * dt
    o.set_limiter(True)
    U = U0
    for s in range(2):
        U, r = o.step(U, po.RK_TVD3, s * dt, dt)
    assert np.abs(U2 - U).max() / np.abs(U).max() < 1e-12
    assert abs(res[1] - r) <= 1e-10 * abs(r)
    assert np.array_equal(O, o.sensor(U2, (0.05, 0.02, 0.01, 0.005), (5e-5, 2e-5, 1e-5, 5e-6), 1, 4))
    out, i = [], 0
    for p0, p1 in zip(O0, O):
        b = U2[i:i + 4 * (p0 + 1) * (p0 + 2) // 2].reshape(4, -1)
        i += b.size
        c = np.zeros((4, (p1 + 1) * (p1 + 2) // 2))
        k = min(b.shape[1], c.shape[1])
        c[:, :k] = b[:, :k]
        out.append(c.ravel())
    assert np.array_equal(U3, np.concatenate(out))


@pytest.mark.skipif(not os.path.exists(EXE), reason="C++ driver not built")
def test_cpp_dropin_maxwell_tets():
    import dgpar
    print("\n[gpu case] C++ drop-in Maxwell", flush=True)
    pre, dt, res = run_exe("maxwell")
    U0, R, U2 = (np.fromfile(pre + x) for x in (".U0", ".R", ".U2"))
    m = dgpar.mesh_tet_structured(2, 2, 2)
    o = po.Oracle(3, m.vids, m.xyz, 3, po.LAW_MAXWELL3D, po.FLUX_UPWIND)
    Uo = o.project(3)
    assert np.abs(U0 - Uo).max() / np.abs(Uo).max() < 1e-13
    Ro, _ = o.residual(U0)
    assert np.abs(R - Ro).max() / np.abs(Ro).max() < 1e-12
    U = U0
    for s in range(2):
        U, r = o.step(U, po.RK4, s * dt, dt)
    assert np.abs(U2 - U).max() / np.abs(U).max() < 1e-12


@pytest.mark.skipif(not os.path.exists(EXE), reason="C++ driver not built")
def test_cpp_dropin_write_read_export():
    print("\n[gpu case] C++ drop-in checkpoint, plot and restart", flush=True)
    d = tempfile.mkdtemp()
    pre = os.path.join(d, "run")
    r = subprocess.run([EXE, pre, "restart"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = dict(l.split(None, 1) for l in r.stdout.splitlines())
    dt = float(lines["dt"])
    t, it = lines["restart"].split()
    assert float(t) == 2 * dt and int(it) == 2
    Ua, Ub = np.fromfile(pre + ".U4a"), np.fromfile(pre + ".U4b")
    assert Ua.size == 200 * 40 and (Ua == Ub).all()
    for f in ("velocity", "pressure", "mach", "density", "error", "entropy", "limit", "order"):
        with open("%s-sample2%s.pos" % (pre, f)) as fh:
            body = fh.read().splitlines()
        assert body[0] == 'View "Exported field" {' and len(body) == 2 + 200 * 9


@pytest.mark.skipif(not os.path.exists(EXE), reason="C++ driver not built")
def test_cpp_dropin_scalarpb_with_a_velocity_field():
    """the reference's ScalarPb written as its scenario is: ScalarLaw(&i, &v, &i)
    with the rotating VelocityOne FieldEvaluator and CaseOne's bump as the far
    field the constructor projects, P1 (SSP-RK2, the integrator the reference
    picks for order 1)"""
    print("\n[gpu case] C++ drop-in ScalarPb", flush=True)
    pre, dt, res = run_exe("scalarpb")
    U0, R, U2 = np.fromfile(pre + ".U0"), np.fromfile(pre + ".R"), np.fromfile(pre + ".U2")
    vids, xyz, bc = mesh_tri(12, 12, x0=-1.0, y0=-1.0, lx=2.0, ly=2.0, jitter=0.1, seed=31)
    o = po.Oracle(2, vids, xyz, 1, po.LAW_SCALAR, po.FLUX_UPWIND, farfield=(0.0,))
    o.set_velocity_field(lambda x, y: (2 * np.pi * y, -2 * np.pi * x))

    def bump(x, y, z, t):
        r = np.hypot(x + 0.5, y)
        return (np.cos(2 * np.pi * r) ** 2 if r <= 0.25 else 0.0,)
    UP, UPo = np.fromfile(pre + ".UP"), o.project_fn(bump)
    assert np.abs(UPo).max() > 0.1 and np.abs(UP - UPo).max() <= 1e-13 * np.abs(UPo).max()
    Ro, _ = o.residual(U0)
    assert np.abs(R - Ro).max() / np.abs(Ro).max() < 1e-12
    assert abs(dt - o.dt(U0, 0.1)) <= 1e-14 * dt
    U = U0
    for s in range(2):
        U, r = o.step(U, po.RK_TVD2, s * dt, dt)
    assert np.abs(U2 - U).max() / np.abs(U).max() < 1e-12
    assert abs(res[1] - r) <= 1e-10 * abs(r)


@pytest.mark.skipif(not os.path.exists(EXE), reason="C++ driver not built")
def test_cpp_dropin_local_time_stepping():
    print("\n[gpu case] C++ drop-in local time stepping", flush=True)
    from helpers import graded_nodes, mesh_tri_graded
    pre, dt, res = run_exe("lts")
    U0, L, U2 = (np.fromfile(pre + x) for x in (".U0", ".L", ".U2"))
    xs = graded_nodes(12, 0.0, 10.0, (4, 2, 1, 1, 2, 4))
    vids, xyz = mesh_tri_graded(xs, xs)
    o = po.Oracle(2, vids, xyz, 1, po.LAW_EULER2D, po.FLUX_LF, gamma=1.4, farfield=(1.0, 1.0, 1.0, 3.5))
    lev, n = o.lts_levels(3)
    np.testing.assert_array_equal(L.astype(np.int32), lev)
    assert abs(dt - o.lts_dt(U0, 0.1, lev)) <= 1e-14 * dt
    assert res == [0.0, 0.0]  # RungeKuttaTVD2adaptive::advanceInTime returns 0
    U = U0
    for _ in range(2):
        U = o.lts_step(U, lev, dt)
    assert np.abs(U2 - U).max() / np.abs(U).max() < 1e-12


@pytest.mark.skipif(not os.path.exists(EXE), reason="C++ driver not built")
def test_cpp_dropin_h_adaptation():
    print("\n[gpu case] C++ drop-in h-adaptation", flush=True)
    from helpers import refine_tri
    pre, dt, res = run_exe("adapt")
    M, U0, U2 = (np.fromfile(pre + x) for x in (".M", ".U0", ".U2"))
    vids, xyz, bc = mesh_tri(10, 10, lx=10.0, ly=10.0, jitter=0.1, seed=77)
    c = xyz.mean(axis=1)
    v, x, b, par = refine_tri(vids, xyz, bc, np.hypot(c[:, 0] - 5.0, c[:, 1] - 5.0) < 2.0)
    np.testing.assert_allclose(M.reshape(-1, 3, 2), x, rtol=0, atol=1e-15)
    o = po.Oracle(2, v, x, 2, po.LAW_EULER2D, po.FLUX_LF, gamma=1.4, farfield=(1.0, 1.0, 1.0, 3.5))
    assert abs(dt - o.dt(U0, 0.1)) <= 1e-14 * dt
    U = U0
    for s in range(2):
        U, _ = o.step(U, po.RK_TVD3, s * dt, dt)
    assert np.abs(U2 - U).max() / np.abs(U).max() < 1e-12
