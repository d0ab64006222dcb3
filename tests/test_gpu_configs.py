"""GPU: the five benchmark configurations at their full sizes, through the C ABI,
against the oracle and the exact solutions.

- C1  scalar advection, 32x32x2 periodic unit square, P2, upwind, Gaussian
      pulse, 100 SSP-RK3 steps at CFL 0.1 (FP64): every step against the oracle
      (1e-12), and the L2 error against the translated pulse bounded by the
      projection error of the exact field;
- C2  isentropic vortex, 256x256x2 jittered periodic [0,10]^2, P3, LF, RK4:
      residual and 5 RK4 steps against the oracle (1e-12); after 500 steps
      the FP32 form against the FP64 form (density L2 within 1e-5) and each
      form's per-field L2 error against the exact vortex;
- C3  Sod tube, 512x128x2 jittered, walls, seeded orders P1-P4, Roe with the
      entropy fix, limiter and sensor: one limited SSP-RK3 step against the
      oracle (1e-12), limiter flags and sensor orders bit for bit on the same
      state; then the whole run to t = 0.2 (sensor every 20 steps) against
      Toro's exact solution, with mass and energy conserved;
- C4  Maxwell TE(1,1,1) cavity, 48^3 x 6 Kuhn tetrahedra, P3, upwind, RK4:
      residual against the oracle (1e-12); 200 steps, L2 error against the
      cavity mode and the FP64 energy drift (non-increasing, small);
- C5  adjoint of the 128x128x2 vortex, P3, LF, 200 RK4 steps, a checkpoint
      every 20: the gradient of the density misfit against central finite
      differences of the FP64 forward on 16 seeded directions (h = 1e-6) and
      against the tangent-linear model of the 200 steps (dot-product test).
Each test prints the figures it checks."""
import math
import os
import sys

import numpy as np
import pytest

from helpers import VORTEX
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

SEED = 20240611
SOD_FF = (1.0, 0.0, 0.0, 2.5, 0.125, 0.0, 0.0, 0.25)
S_HI = (0.05, 0.02, 0.01, 0.005)
S_LO = (5e-5, 2e-5, 1e-5, 5e-6)
THREADS = max(1, min(16, len(os.sched_getaffinity(0))))


def announce(name):
    print("\n[gpu case] %s" % name, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def test_c1_scalar_gaussian_100_tvd3_steps():
    import dgpar
    announce("C1 scalar P2 32x32x2, 100 SSP-RK3 steps")
    m = dgpar.mesh_tri_structured(32, 32)
    lw = dgpar.Law(dgpar.LAW_SCALAR, dgpar.FLUX_UPWIND, vel=(1.0, 0.5), farfield=(0.0,))
    o = po.Oracle(2, m.vids, m.xyz, 2, po.LAW_SCALAR, po.FLUX_UPWIND, vel=(1.0, 0.5), farfield=(0.0,))
    with dgpar.DGAnalysis(m, lw, 2) as g:
        g.project(dgpar.FIELD_GAUSSIAN)
        U = o.project(0)
        assert rel(g.get_state(), U) < 1e-14
        dt = o.dt(U, 0.1)
        assert abs(g.computeDt(0.1) - dt) <= 1e-15 * dt
        t = 0.0
        for _ in range(100):
            U, ro = o.step(U, po.RK_TVD3, t, dt)
            t += dt
        r = g.advanceInTime(0.0, dt, scheme=dgpar.RK_TVD3, nsteps=100)
        err_state = rel(g.get_state(), U)
        eg = g.l2_error(dgpar.FIELD_GAUSSIAN, (), t)[0]
    eo = o.l2_error(U, 0, (), t)[0]
    # the L2 projection is the best approximation in the element space
    ep = o.l2_error(o.project(0, (), t), 0, (), t)[0]
    print("C1: t=%.6f state rel %.2e  L2 err GPU %.4e oracle %.4e projection %.4e" % (t, err_state, eg, eo, ep))
    assert err_state < 1e-12
    assert abs(r - ro) <= 1e-10 * abs(ro)
    assert abs(eg - eo) <= 1e-10 * eo
    assert 0.999 * ep <= eg <= 3.0 * ep


def vortex_mesh(n):
    import dgpar
    return dgpar.mesh_tri_structured(n, n, lx=10.0, ly=10.0, jitter=0.1, seed=SEED)


def test_c2_vortex_residual_and_rk4_steps_vs_oracle():
    import dgpar
    announce("C2 vortex P3 256x256x2: residual and 5 RK4 steps vs oracle")
    m = vortex_mesh(256)
    o = po.Oracle(2, m.vids, m.xyz, 3, po.LAW_EULER2D, po.FLUX_LF, gamma=1.4)
    with dgpar.DGAnalysis(m, dgpar.EulerLaw(dgpar.FLUX_LF, 1.4), 3) as g:
        g.project(dgpar.FIELD_VORTEX, VORTEX)
        U = o.project(1, VORTEX)
        assert rel(g.get_state(), U) < 1e-14
        # The residual is a sum of volume and face terms; on a smooth field it is
        # far smaller than its terms (cancellation, about 1/h = 26 here), so both
        # sides' FP64 rounding of the terms, summed in their own orders (and with
        # the device's fused multiply-adds), is measured against the magnitude of
        # the terms: per DOF |R_gpu - R_oracle| / sum |terms|, the relative error
        # of the summation.  Against max |R| the smooth vortex shows 9e-11.
        def scaled(Ug):
            Rg = g.residual()
            Ro = o.residual(Ug, nthreads=THREADS)[0]
            S = o.residual_scale(Ug, nthreads=THREADS)
            return float((np.abs(Rg - Ro) / np.maximum(S, 1e-300)).max()), rel(Rg, Ro)
        er_smooth, er_smooth_maxr = scaled(U)
        # the vortex with 0.2% of the field scale added to every mode above the mean
        V = U + 0.002 * np.abs(U).max() * np.random.default_rng(5).standard_normal(U.size) * (
            np.tile(np.repeat([1.0, 0.2, 0.2, 1.0], 10), m.n_elems) * (np.arange(U.size) % 10 > 0))
        g.set_state(V)
        er, er_maxr = scaled(V)
        g.set_state(U)
        dt = o.dt(U, 0.1)
        assert abs(g.computeDt(0.1) - dt) <= 1e-14 * dt
        t = 0.0
        for _ in range(5):
            U, ro = o.step(U, po.RK4, t, dt, nthreads=THREADS)
            t += dt
        r = g.advanceInTime(0.0, dt, scheme=dgpar.RK4, nsteps=5)
        es = rel(g.get_state(), U)
    print("C2: residual vs the terms' magnitude: smooth vortex %.2e, perturbed %.2e (vs max|R|: %.2e, %.2e); "
          "state after 5 RK4 steps rel %.2e" % (er_smooth, er, er_smooth_maxr, er_maxr, es))
    assert er_smooth < 1e-12
    assert er < 1e-12
    assert es < 1e-12
    assert abs(r - ro) <= 1e-10 * abs(ro)


def test_c2_fp32_vs_fp64_500_steps_and_exact_errors():
    import dgpar
    announce("C2 vortex P3 256x256x2: 500 RK4 steps FP32 vs FP64")
    m = vortex_mesh(256)
    res = {}
    for prec in (dgpar.PREC_FP64, dgpar.PREC_FP32):
        with dgpar.DGAnalysis(m, dgpar.EulerLaw(dgpar.FLUX_LF, 1.4), 3, precision=prec) as g:
            g.project(dgpar.FIELD_VORTEX, VORTEX)
            dt = g.computeDt(0.1)
            g.time_advance(dgpar.RK4, dt, 500)
            res[prec] = (g.get_state(), g.l2_error(dgpar.FIELD_VORTEX, VORTEX, 500 * dt))
    X = m.xyz
    detj = np.abs((X[:, 1, 0] - X[:, 0, 0]) * (X[:, 2, 1] - X[:, 0, 1]) -
                  (X[:, 2, 0] - X[:, 0, 0]) * (X[:, 1, 1] - X[:, 0, 1]))
    d = (res[1][0] - res[0][0]).reshape(m.n_elems, 4, 10)
    # orthonormal modes: int (u - v)^2 over an element = |J| sum_k (u_k - v_k)^2
    l2 = np.sqrt((d ** 2 * detj[:, None, None]).sum(axis=(0, 2)))
    print("C2: per-field L2 FP32 - FP64 %s; vs exact FP64 %s FP32 %s"
          % (np.array2string(l2, precision=3), np.array2string(res[0][1], precision=3),
             np.array2string(res[1][1], precision=3)))
    assert l2[0] <= 1e-5
    assert (l2 <= 1e-4).all()
    # the FP32 form's error against the exact vortex stays at the FP64 form's
    # error plus its own rounding difference (triangle inequality)
    assert (res[1][1] <= res[0][1] + l2 * (1 + 1e-9)).all()


def tube_mesh():
    import dgpar
    m0 = dgpar.mesh_tri_structured(512, 128, lx=1.0, ly=0.25, jitter=0.1, seed=SEED, periodic_x=False,
                                   periodic_y=False,
                                   side_tags=(dgpar.TAG_WALL, dgpar.TAG_FARFIELD2, dgpar.TAG_WALL, dgpar.TAG_FARFIELD))
    orders = np.random.default_rng(SEED).integers(1, 5, m0.n_elems).astype(np.int32)
    return m0, orders


def tube_law():
    import dgpar
    return dgpar.EulerLaw(dgpar.FLUX_ROE_EF, 1.4, farfield=SOD_FF)


def test_c3_tube_limited_step_flags_and_orders_vs_oracle():
    import dgpar
    announce("C3 Sod tube 512x128x2 mixed P1-P4: limited SSP-RK3 step, flags, orders")
    m0, orders = tube_mesh()
    mesh = dgpar.Mesh(2, m0.vids, m0.xyz, m0.bc, order=orders)
    o = po.Oracle(2, m0.vids, m0.xyz, orders, po.LAW_EULER2D, po.FLUX_ROE_EF, gamma=1.4, farfield=SOD_FF, bc=m0.bc)
    with dgpar.DGAnalysis(mesh, tube_law(), 1) as g:
        g.project(dgpar.FIELD_SOD, (0.5,))
        U0 = o.project(2, (0.5,))
        assert rel(g.get_state(), U0) < 1e-13
        g.set_limiter(True, 0.0, 1.0)
        f0 = g.limit()
        U, fo = o.limit(U0)
        assert np.array_equal(f0, fo) and fo.any()
        assert rel(g.get_state(), U) < 1e-13
        dt = o.dt(U, 0.1)
        assert abs(g.computeDt(0.1) - dt) <= 1e-13 * dt
        o.set_limiter(True)
        Uo, ro = o.step(U, po.RK_TVD3, 0.0, dt, nthreads=THREADS)
        # the bucket-by-bucket launches (schedule shape 1, compiled apart from the
        # persistent launch) give its state to FP64 rounding
        g.set_schedule(shape=1)
        r1 = g.advanceInTime(0.0, dt, scheme=dgpar.RK_TVD3, nsteps=1)
        U1 = g.get_state()
        g.set_state(U)
        g.set_schedule(shape=0)
        r = g.advanceInTime(0.0, dt, scheme=dgpar.RK_TVD3, nsteps=1)
        assert rel(U1, g.get_state()) < 1e-12 and abs(r1 - r) <= 1e-12 * abs(r)
        es = rel(g.get_state(), Uo)
        # limiter flags and sensor orders on one state, bit for bit
        Ug = g.get_state()
        fg = g.limit()
        Ul, fl = o.limit(Ug)
        el = rel(g.get_state(), Ul)
        Ug = g.get_state()
        want = o.sensor(Ug, S_HI, S_LO, 1, 4)
        got = g.select_orders(S_HI, S_LO, 1, 4, apply=False)
    print("C3: step rel %.2e, limited elements %d (after the step %d), limit rel %.2e, sensor changes %d"
          % (es, int((fo > 0).sum()), int((fl > 0).sum()), el, int((want != orders).sum())))
    assert es < 1e-12
    assert abs(r - ro) <= 1e-10 * abs(ro)
    assert np.array_equal(fg, fl)
    assert el < 1e-13
    assert np.array_equal(got, want)
    assert (want != orders).any()


def test_c3_tube_to_t02_against_exact_sod():
    import dgpar
    announce("C3 Sod tube 512x128x2 mixed P1-P4 to t = 0.2 vs the exact solution")
    m0, orders = tube_mesh()
    mesh = dgpar.Mesh(2, m0.vids, m0.xyz, m0.bc, order=orders)
    T = 0.2
    with dgpar.DGAnalysis(mesh, tube_law(), 1) as g:
        g.project(dgpar.FIELD_SOD, (0.5,))
        g.set_limiter(True, 0.0, 1.0)
        g.limit()
        n0 = g.norms()
        dt = g.computeDt(0.1)
        nsteps = int(math.ceil(T / dt - 1e-9))
        t, done, passes = 0.0, 0, 0
        while done < nsteps:
            n = min(20, nsteps - done)
            h = [dt] * n
            if done + n == nsteps:
                h[-1] = T - (t + (n - 1) * dt)  # the last step lands on T
            if n > 1:
                g.advanceInTime(t, dt, scheme=dgpar.RK_TVD3, nsteps=n - 1, want_resid=False)
            g.advanceInTime(t + (n - 1) * dt, h[-1], scheme=dgpar.RK_TVD3, nsteps=1, want_resid=False)
            t += (n - 1) * dt + h[-1]
            done += n
            if done < nsteps:
                g.select_orders(S_HI, S_LO, 1, 4, apply=True)
                passes += 1
        n1 = g.norms(T, dgpar.FIELD_SOD, (0.5,))
        hist = np.bincount(g.orders(), minlength=5)[1:]
        U = g.get_state()
    area = 0.25
    l1 = n1["l1"][0] / area
    l2 = n1["l2"][0] / math.sqrt(area)
    dm = abs(n1["totals"][0] - n0["totals"][0]) / n0["totals"][0]
    de = abs(n1["totals"][3] - n0["totals"][3]) / n0["totals"][3]
    print("C3: %d steps (dt %.4e), %d sensor passes, orders P1-P4 %s; density L1 %.4e L2 %.4e (per unit area); "
          "mass drift %.2e energy drift %.2e" % (nsteps, dt, passes, hist.tolist(), l1, l2, dm, de))
    assert abs(t - T) <= 1e-14
    assert np.isfinite(U).all()
    # no wave reaches the ends by t = 0.2 (head at x = 0.263, shock at 0.850),
    # so the far-field ends carry no mass or energy: both totals are conserved
    assert dm <= 1e-12 and de <= 1e-12
    # the initial step is 0.08 (L1 per unit area) away from the exact solution at
    # t = 0.2; the scheme must resolve the waves to within a few cells
    assert l1 <= 5e-3


def test_c4_maxwell_tets_residual_and_200_steps():
    import dgpar
    announce("C4 Maxwell cavity 48^3 x 6 tets P3: residual vs oracle, 200 RK4 steps")
    m = dgpar.mesh_tet_structured(48, 48, 48)
    with dgpar.DGAnalysis(m, dgpar.MaxwellLaw(dgpar.FLUX_UPWIND), 3) as g:
        o = po.Oracle(3, m.vids, m.xyz, 3, po.LAW_MAXWELL3D, po.FLUX_UPWIND)
        U = o.project(3)
        g.project(dgpar.FIELD_CAVITY)
        ep = rel(g.get_state(), U)
        V = U + 0.01 * np.random.default_rng(4).standard_normal(U.size)
        g.set_state(V)
        er = rel(g.residual(), o.residual(V, nthreads=THREADS)[0])
        dt = o.dt(U, 0.1)
        del o, V
        assert abs(g.computeDt(0.1) - dt) <= 1e-15 * dt
        g.project(dgpar.FIELD_CAVITY)
        e0 = g.norms()["energy"].sum()
        ep0 = g.l2_error(dgpar.FIELD_CAVITY, (), 0.0)
        g.time_advance(dgpar.RK4, dt, 200)
        T = 200 * dt
        n1 = g.norms(T, dgpar.FIELD_CAVITY)
        e1 = n1["energy"].sum()
    drift = (e1 - e0) / e0
    # the FP32 form: the same 200 steps, energy from the FP64 step-start state
    with dgpar.DGAnalysis(m, dgpar.MaxwellLaw(dgpar.FLUX_UPWIND), 3, precision=dgpar.PREC_FP32) as h:
        h.project(dgpar.FIELD_CAVITY)
        f0 = h.norms()["energy"].sum()
        h.time_advance(dgpar.RK4, dt, 200)
        m1 = h.norms(T, dgpar.FIELD_CAVITY)
        drift32 = (m1["energy"].sum() - f0) / f0
    print("C4: projection rel %.2e, residual rel %.2e; t=%.4f L2 error E,H %s (t=0: %s); energy drift %.3e; "
          "FP32 form: L2 error %s, energy drift %.3e"
          % (ep, er, T, np.array2string(n1["l2"], precision=3), np.array2string(ep0, precision=3), drift,
             np.array2string(m1["l2"], precision=3), drift32))
    # FP32 rounding in the fields stays below the discretisation error's size
    # scale (the cavity mode's amplitude is 1); the energy drift stays small
    assert (m1["l2"] <= 1e-5).all()
    assert abs(drift32) <= 1e-5
    assert ep < 1e-13
    assert er < 1e-12
    # the upwind flux only dissipates: energy never grows; the drift is of the
    # size of the discretisation error
    assert -1e-6 <= drift <= 1e-13
    # P3 on h = 1/48: the error stays at the level of the initial projection's
    assert (n1["l2"] <= 1e-6).all()
    assert n1["l2"].max() <= 5 * ep0.max()


def test_c5_adjoint_gradient_fd_16_directions_and_tlm():
    import dgpar
    announce("C5 adjoint vortex 128x128x2 P3, 200 RK4 steps, checkpoint every 20")
    m = vortex_mesh(128)
    shifted = (5.0, 5.5, 5.3, 1.0, 1.0, 10.0)
    nsteps, every, h = 200, 20, 1e-6
    o = po.Oracle(2, m.vids, m.xyz, 3, po.LAW_EULER2D, po.FLUX_LF, gamma=1.4)
    with dgpar.DGAnalysis(m, dgpar.EulerLaw(dgpar.FLUX_LF, 1.4), 3) as g:
        g.project(dgpar.FIELD_VORTEX, VORTEX)
        U0 = g.get_state()
        dt = g.computeDt(0.1)
        g.project(dgpar.FIELD_VORTEX, shifted)
        g.time_advance(dgpar.RK4, dt, nsteps)
        target = g.get_state()
        g.set_state(U0)
        J = g.adjoint_forward(dt, nsteps, every, target)
        grad = g.adjoint_backward()

        def J_of(U):
            g.set_state(U)
            return g.adjoint_forward(dt, nsteps, every, target)

        worst = 0.0
        for i in range(16):
            V = np.random.default_rng(SEED + i).standard_normal(U0.size) * (0.05 * np.abs(U0).max())
            fd = (J_of(U0 + h * V) - J_of(U0 - h * V)) / (2 * h)
            gv = float(np.dot(grad, V))
            e = abs(gv - fd) / abs(fd)
            worst = max(worst, e)
            print("C5: direction %2d  grad.V %.10e  FD %.10e  rel %.2e" % (i, gv, fd, e), flush=True)
        # tangent-linear model of the 200 steps: <S, T V> = <grad_S, V>
        S = np.random.default_rng(SEED + 99).standard_normal(U0.size)
        V = np.random.default_rng(SEED + 98).standard_normal(U0.size) * (0.05 * np.abs(U0).max())
        g.set_state(U0)
        g.adjoint_forward(dt, nsteps, every)
        gS = g.adjoint_backward(S)
        detj = o.detj()
        dtj = np.repeat(dt / detj, U0.size // detj.size)
        U, dU = U0.copy(), V.copy()
        for _ in range(nsteps):
            k, dk = [], []
            Us, dUs = U, dU
            for c in (0.5, 0.5, 1.0, None):
                g.set_state(Us)
                k.append(dtj * g.residual())
                dk.append(dtj * g.residual_jvp(Us, dUs))
                if c is not None:
                    Us, dUs = U + c * k[-1], dU + c * dk[-1]
            U = U + (k[0] + 2 * k[1] + 2 * k[2] + k[3]) / 6
            dU = dU + (dk[0] + 2 * dk[1] + 2 * dk[2] + dk[3]) / 6
        lhs, rhs = float(np.dot(S, dU)), float(np.dot(gS, V))
        fwd_ms, bwd_ms = g.adjoint_timing()
    tlm = abs(lhs - rhs) / np.abs(S * dU).sum()
    print("C5: J %.10e, worst FD rel %.2e, TLM identity rel %.2e, backward/forward %.2f"
          % (J, worst, tlm, bwd_ms / fwd_ms))
    assert worst <= 1e-6
    assert tlm <= 1e-11
