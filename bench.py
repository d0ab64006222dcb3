#!/usr/bin/env python3
"""Benchmark: 2D Euler isentropic vortex, P3, 256x256x2 = 131072 jittered periodic
triangles on [0,10]^2, local Lax-Friedrichs flux, RK4 at CFL 0.1.

One "step" is one RK4 step (4 stages, each a fused residual + stage-update
kernel) over the whole mesh, replayed from a hipGraph; inputs are resident in
HBM before the timed region.  The headline value is the reduced form (FP32
volume/face arithmetic, FP64 stage accumulation and inverse mass); the plain
FP64 form is timed beside it, with the density L2 error of both against the
exact vortex and against each other.  The discrete adjoint is timed on the
128x128x2 vortex mesh: 200 RK4 steps forward with a checkpoint every 20 steps,
then the backward sweep (segment recomputation + fused reverse stages) for the
gradient of the density misfit against the solution started from the vortex
shifted by (0.5, 0.3); the ratio backward/forward is reported.  The CPU
baseline is the reference's own serial code (its DGAnalysis and TimeIntegrator,
built from its sources into oracle/_ref) running RK4 steps on the same mesh;
the FP64 CPU oracle (a restatement of the reference algorithm) on a bounded
sample on this host's cores is reported beside it.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "dg-parallel_amd"))

NX = 256
ORDER = 3
CFL = 0.1
SEED = 20240611
VORTEX = (5.0, 5.0, 5.0, 1.0, 1.0, 10.0)
HBM_PEAK_GBS = 8000.0


def bytes_per_dof_stage(fp32, nf, np_):
    """Algorithmic HBM bytes per DOF and RK4 stage of the fused stage kernel:
    stage input read once, FP64 step-start state and accumulator, stage output,
    plus the element table (|J| J^-1: 4 doubles, |J|, 3 int32 links) per element.
    The FP32 form's first stage reads U^n rounded to FP32 (written by the step
    before's last stage) as its input and U^n in FP64 for the update."""
    s = 4 if fp32 else 8
    st = [(4 + 8 if fp32 else 8) + 8 + s,  # stage 1: read U^n (FP32 copy, FP64), write acc, write U1
          s + 8 + 16 + s,                   # stage 2: read U1, U^n, acc; write acc, U2
          s + 8 + 16 + s,                   # stage 3
          s + 8 + 8 + 8 + (4 if fp32 else 0)]  # stage 4: read U3, U^n, acc; write U^{n+1} (and its FP32 copy)
    table = (4 * 8 + 8 + 3 * 4) / (nf * np_)
    return sum(st) / 4.0 + table


def measured_traffic(form):
    """HBM bytes per launch of the stage kernel for this workload, from the two
    rocprofv3 counter passes (FETCH_SIZE, WRITE_SIZE, each scaled by what the
    same counter reads for a 512 MiB device copy) summarised under profiles/."""
    path = os.path.join(ROOT, "profiles", "hbm_counters_tri_stage_tiled_euler_vortex_p3.md")
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch_%s:" % form):
                    return int(line.split(":")[1])
    except OSError:
        pass
    return None


def tuned(g, scheme, dt, tune):
    """the tuner's choice of launch schedule (timed on the leg's own state,
    which it restores), or the default schedule"""
    if not tune:
        return {"best": g.schedule(), "candidates": []}
    rep = g.autotune(scheme, dt, steps=20)
    return {"best": rep["best"], "candidates": [[c[0], c[1], round(c[2], 4)] for c in rep["candidates"]],
            "fields": "shape (0: 4 waves per 64-element tile, 1: 8 waves), grid (persistent workgroups, 0: default), "
                      "ms per step"}


def run_form(dgpar, mesh, precision, steps, warmup, kernel=0, tune=True):
    law = dgpar.EulerLaw(dgpar.FLUX_LF, 1.4)
    g = dgpar.DGAnalysis(mesh, law, ORDER, precision=precision, use_graph=True, device=0, kernel=kernel)
    g.project(dgpar.FIELD_VORTEX, VORTEX)
    dt = g.computeDt(CFL)
    sched = tuned(g, dgpar.RK4, dt, tune)
    if warmup > 0:
        g.time_advance(dgpar.RK4, dt, warmup)
    g.project(dgpar.FIELD_VORTEX, VORTEX)
    # bracket the timed steps with device-wide synchronisation on the library's
    # HIP runtime (one process, one GPU: no barrier needed)
    dgpar.device_synchronize()
    t0 = time.perf_counter()
    ms = g.time_advance(dgpar.RK4, dt, steps)
    dgpar.device_synchronize()
    wall = time.perf_counter() - t0
    err = g.l2_error(dgpar.FIELD_VORTEX, VORTEX, steps * dt)
    state = g.get_state()
    ndof = g.ndof
    g.close()
    return dict(ms=ms, wall=wall, dt=dt, err=err, state=state, ndof=ndof, schedule=sched)


def run_adjoint(dgpar, steps, every):
    mesh = dgpar.mesh_tri_structured(128, 128, lx=10.0, ly=10.0, jitter=0.1, seed=SEED)
    law = dgpar.EulerLaw(dgpar.FLUX_LF, 1.4)
    g = dgpar.DGAnalysis(mesh, law, ORDER, precision=dgpar.PREC_FP64, use_graph=True, device=0)
    shifted = (5.0, 5.5, 5.3, 1.0, 1.0, 10.0)
    g.project(dgpar.FIELD_VORTEX, VORTEX)
    dt = g.computeDt(CFL)
    g.project(dgpar.FIELD_VORTEX, shifted)
    g.time_advance(dgpar.RK4, dt, steps)
    target = g.get_state()
    # warm-up pass (module loads, buffer allocation), then the timed pair
    g.project(dgpar.FIELD_VORTEX, VORTEX)
    g.adjoint_forward(dt, 2 * every, every, target)
    g.adjoint_backward()
    g.project(dgpar.FIELD_VORTEX, VORTEX)
    fwd_plain_ms = g.time_advance(dgpar.RK4, dt, steps)
    g.project(dgpar.FIELD_VORTEX, VORTEX)
    J = g.adjoint_forward(dt, steps, every, target)
    grad = g.adjoint_backward()
    fwd_ms, bwd_ms = g.adjoint_timing()
    ndof = g.ndof
    g.close()
    return {"workload": "euler2d-isentropic-vortex P3 LF RK4, 128x128x2 periodic [0,10]^2 jitter 0.1h seed %d" % SEED,
            "steps": steps, "ckpt_every": every, "dof": ndof, "objective_J": J,
            "grad_norm": float(np.sqrt(np.dot(grad, grad))),
            "fwd_ms": fwd_ms, "bwd_ms": bwd_ms, "fwd_plain_graph_ms": fwd_plain_ms,
            "ratio_bwd_over_fwd": bwd_ms / fwd_ms, "ratio_bwd_over_fwd_plain": bwd_ms / fwd_plain_ms,
            "fwd_value": ndof * 4 * steps / (fwd_ms / 1e3), "unit": "DOF*stage/s"}


def run_scalar(dgpar, steps=100, tune=True):
    """Scalar advection, velocity (1, 0.5), periodic unit square as 32x32 squares
    split into 2048 triangles, P2, upwind flux, Gaussian pulse, SSP-RK3 at CFL
    0.1, 100 steps, FP64; the L2 error against the translated pulse after the
    timed steps.  12 288 DOF: the step is bound by launch latency, not HBM."""
    mesh = dgpar.mesh_tri_structured(32, 32)
    law = dgpar.Law(dgpar.LAW_SCALAR, dgpar.FLUX_UPWIND, vel=(1.0, 0.5), farfield=(0.0,))
    g = dgpar.DGAnalysis(mesh, law, 2, precision=dgpar.PREC_FP64, use_graph=True, device=0)
    g.project(dgpar.FIELD_GAUSSIAN)
    dt = g.computeDt(CFL)
    sched = tuned(g, dgpar.RK_TVD3, dt, tune)
    g.time_advance(dgpar.RK_TVD3, dt, 10)
    g.project(dgpar.FIELD_GAUSSIAN)
    dgpar.device_synchronize()
    ms = g.time_advance(dgpar.RK_TVD3, dt, steps)
    err = g.l2_error(dgpar.FIELD_GAUSSIAN, (), steps * dt)
    ndof = g.ndof
    g.close()
    return {"workload": "scalar advection v=(1,0.5), 32x32x2 periodic unit square, P2, upwind, Gaussian pulse, "
                        "SSP-RK3 CFL0.1", "dof": ndof, "steps": steps, "dt": dt, "ms_per_step": ms / steps,
            "value": ndof * 3 * steps / (ms / 1e3), "unit": "DOF*stage/s", "l2_error_vs_exact": float(err[0]),
            "schedule": sched}


def run_mixed(dgpar, steps, sense_every=20, tune=True, t_end=0.2):
    """Shock tube with mixed orders: 512x128x2 jittered triangles on [0,1]x[0,0.25],
    walls at the bottom and top, the two initial states as far field at the ends,
    a seeded order map uniform on {1,2,3,4}, Roe flux with entropy fix, the
    triangle limiter after every SSP-RK3 stage, and the order sensor with
    re-bucketing every sense_every steps.  steps < 0: the whole run to t_end
    (the last step lands on it), with the density error against the exact Sod
    solution.  The timed region holds the steps (one persistent launch per stage
    + the limiter) and the sensor passes between them; both are also reported
    apart."""
    nx, ny = 512, 128
    mesh0 = dgpar.mesh_tri_structured(nx, ny, lx=1.0, ly=0.25, jitter=0.1, seed=SEED, periodic_x=False,
                                      periodic_y=False, side_tags=(dgpar.TAG_WALL, dgpar.TAG_FARFIELD2,
                                                                   dgpar.TAG_WALL, dgpar.TAG_FARFIELD))
    E = mesh0.n_elems
    orders = np.random.default_rng(SEED).integers(1, 5, E).astype(np.int32)
    mesh = dgpar.Mesh(2, mesh0.vids, mesh0.xyz, mesh0.bc, order=orders)
    law = dgpar.EulerLaw(dgpar.FLUX_ROE_EF, 1.4, farfield=(1.0, 0.0, 0.0, 2.5, 0.125, 0.0, 0.0, 0.25))
    g = dgpar.DGAnalysis(mesh, law, 1, precision=dgpar.PREC_FP64, use_graph=True, device=0)
    g.project(dgpar.FIELD_SOD, (0.5,))
    g.set_limiter(True, 0.0, 1.0)
    g.limit()
    dt = g.computeDt(CFL)
    s_hi, s_lo = (0.05, 0.02, 0.01, 0.005), (5e-5, 2e-5, 1e-5, 5e-6)
    sched = tuned(g, dgpar.RK_TVD3, dt, tune)
    g.time_advance(dgpar.RK_TVD3, dt, 2)  # warm-up (graph capture)
    # restart from the initial state after the tuner and the warm-up (neither
    # changes the orders)
    g.project(dgpar.FIELD_SOD, (0.5,))
    g.set_limiter(True, 0.0, 1.0)
    g.limit()
    n0 = g.norms()
    full = steps < 0
    if full:
        steps = int(np.ceil(t_end / dt - 1e-9))
    step_ms, sense_ms, dof_stages = 0.0, 0.0, 0.0
    done, t = 0, 0.0
    dgpar.device_synchronize()
    w0 = time.perf_counter()
    while done < steps:
        n = min(sense_every, steps - done)
        last = done + n == steps
        if full and last:
            h = t_end - (t + (n - 1) * dt)  # the last step lands on t_end
            if n > 1:
                step_ms += g.time_advance(dgpar.RK_TVD3, dt, n - 1)
            step_ms += g.time_advance(dgpar.RK_TVD3, h, 1)
            t += (n - 1) * dt + h
        else:
            step_ms += g.time_advance(dgpar.RK_TVD3, dt, n)
            t += n * dt
        dof_stages += g.ndof * 3 * n
        done += n
        if not last:
            t0 = time.perf_counter()
            g.select_orders(s_hi, s_lo, 1, 4, apply=True)
            sense_ms += (time.perf_counter() - t0) * 1e3
    dgpar.device_synchronize()
    wall_ms = (time.perf_counter() - w0) * 1e3
    flags = g.limiter_flags()  # the last stage's limiter pass (no sensor pass since)
    hist = np.bincount(g.orders(), minlength=5)[1:].tolist()
    exact = None
    if full:
        n1 = g.norms(t, dgpar.FIELD_SOD, (0.5,))
        area = 0.25
        exact = {"t": t, "density_l1_per_area": float(n1["l1"][0] / area),
                 "density_l2_per_sqrt_area": float(n1["l2"][0] / np.sqrt(area)),
                 "density_linf": float(n1["linf"][0]),
                 "mass_drift": float((n1["totals"][0] - n0["totals"][0]) / n0["totals"][0]),
                 "energy_drift": float((n1["totals"][3] - n0["totals"][3]) / n0["totals"][3])}
    # algorithmic bytes per DOF*stage: SSP-RK3 stage traffic (stage 1 reads U^n
    # and writes A: 16 B; stages 2, 3 read the stage input and U^n and write:
    # 24 B each), the element table (|J| J^-1, |J|, 3 links: 52 B per element
    # over Nf * mean Np DOF) and the limiter's read of the stage output (8 B)
    mean_np = sum(c * (p + 1) * (p + 2) // 2 for p, c in zip((1, 2, 3, 4), hist)) / E
    b = (16 + 24 + 24) / 3.0 + 52.0 / (4 * mean_np) + 8.0
    ach = b * dof_stages / (step_ms / 1e3) / 1e9
    passes = -(-steps // sense_every) - 1
    out = {"workload": "euler2d Sod tube 512x128x2 [0,1]x[0,0.25] jitter 0.1h, orders P1-P4 seeded, Roe+entropy fix, "
                       "SSP-RK3 CFL0.1, limiter every stage, sensor every %d steps%s"
                       % (sense_every, ", to t=%g" % t_end if full else ""),
           "steps": steps, "dt": dt, "value": dof_stages / (wall_ms / 1e3), "unit": "DOF*stage/s",
           "value_includes": "steps, limiter and sensor + re-bucketing passes (wall clock of the whole run)",
           "value_steps_only": dof_stages / (step_ms / 1e3),
           "ms_per_step": wall_ms / steps, "ms_per_step_steps_only": step_ms / steps,
           "achieved_GBs_steps_only": ach, "frac_steps_only": ach / HBM_PEAK_GBS, "bytes_per_dof_stage": b,
           "sensor_passes": passes, "sensor_rebucket_ms_per_pass": sense_ms / max(1, passes),
           "orders_histogram_p1_p4": hist, "limited_elements_last_stage": int((flags > 0).sum()),
           "positivity_scaled_last_stage": int(((flags & 16) > 0).sum()),
           "steps_to_t0.2": int(np.ceil(0.2 / dt)), "exact_sod": exact, "schedule": sched}
    g.close()
    return out


def run_maxwell_form(dgpar, mesh, precision, steps, tune):
    g = dgpar.DGAnalysis(mesh, dgpar.MaxwellLaw(dgpar.FLUX_UPWIND), 3, precision=precision, use_graph=True)
    g.project(dgpar.FIELD_CAVITY)
    dt = g.computeDt(CFL)
    sched = tuned(g, dgpar.RK4, dt, tune)
    g.time_advance(dgpar.RK4, dt, 2)
    g.project(dgpar.FIELD_CAVITY)
    e0 = float(g.norms()["energy"].sum())
    dgpar.device_synchronize()
    ms = g.time_advance(dgpar.RK4, dt, steps)
    nrm = g.norms(steps * dt, dgpar.FIELD_CAVITY)
    state = g.get_state()
    ndof = g.ndof
    g.close()
    return dict(ms=ms, ndof=ndof, err=[float(x) for x in nrm["l2"]],
                drift=(float(nrm["energy"].sum()) - e0) / e0, state=state, sched=sched)


def run_maxwell(dgpar, steps, n=48, tune=True):
    """Maxwell on tetrahedra: 48^3 boxes x 6 Kuhn tetrahedra on the unit cube, PEC
    walls, P3, upwind flux, RK4 at CFL 0.1, the (1,1,1) cavity mode, in the FP64
    form and in the FP32 form (FP32 arithmetic and stage copies, FP64 step-start
    state, accumulator and update); per form the L2 error of E and H against the
    exact mode after the timed steps and the drift of the total energy (FP64
    sums of the FP64 state), and the FP32 form's L2 distance to the FP64 form."""
    mesh = dgpar.mesh_tet_structured(n, n, n)
    r64 = run_maxwell_form(dgpar, mesh, dgpar.PREC_FP64, steps, tune)
    r32 = run_maxwell_form(dgpar, mesh, dgpar.PREC_FP32, steps, tune)
    ndof = r64["ndof"]
    # algorithmic bytes per DOF*stage: RK4 state traffic and the element table
    # (|J| J^-1, |J|, 4 links per element over 6 x 20 DOF); FP32 stage copies,
    # the first stage reading U^n's FP32 copy and U^n, the last writing both
    table = (9 * 8 + 8 + 4 * 4) / (6 * 20.0)
    b64 = 34.0 + table
    b32 = (24 + 32 + 32 + 32) / 4.0 + table

    def leg(r, b):
        ach = b * ndof / (r["ms"] / (4 * steps) / 1e3) / 1e9
        return {"ms_per_step": r["ms"] / steps, "value": ndof * 4 * steps / (r["ms"] / 1e3), "achieved_GBs": ach,
                "frac": ach / HBM_PEAK_GBS, "bytes_per_dof_stage": b, "l2_error_E_H": r["err"],
                "energy_drift": r["drift"], "schedule": r["sched"]}
    d = (r32["state"] - r64["state"]).reshape(6 * n ** 3, 6, 20)
    X = mesh.xyz
    J = np.stack([X[:, 1] - X[:, 0], X[:, 2] - X[:, 0], X[:, 3] - X[:, 0]], axis=2)
    detj = np.abs(np.linalg.det(J))
    l2diff = np.sqrt((d ** 2 * detj[:, None, None]).sum(axis=(0, 2)))
    f64, f32 = leg(r64, b64), leg(r32, b32)
    return {"workload": "maxwell3d PEC cavity (1,1,1), %d^3 x 6 Kuhn tets, P3, upwind, RK4 CFL0.1" % n,
            "elements": 6 * n ** 3, "dof": ndof, "steps": steps, "unit": "DOF*stage/s",
            "fp64": f64, "fp32": f32, "ratio_fp32_over_fp64": f32["value"] / f64["value"],
            "l2_fp32_vs_fp64_E_H": [float(x) for x in l2diff],
            # the FP64 form's figures at the top level, as the earlier lines had them
            "ms_per_step": f64["ms_per_step"], "value": f64["value"], "achieved_GBs": f64["achieved_GBs"],
            "frac": f64["frac"], "bytes_per_dof_stage": b64, "l2_error_E_H": f64["l2_error_E_H"],
            "schedule": f64["schedule"]}


REF_DRIVER = os.path.join(os.path.dirname(os.path.abspath(__file__)), "oracle", "_ref", "ref_driver")


def write_ref_lattice(mesh, n, d, U):
    """The structured n x n x 2 mesh as a Gmsh file for the reference's driver (the
    (n+1)^2 lattice nodes, the triangles in the generator's order, 510 / 610 lines
    on the periodic sides) and the state U as its state file, in directory d.
    Returns (mesh path, state path, number of triangles)."""
    N1 = n + 1
    X = mesh.xyz.reshape(n, n, 2, 3, 2)  # [j][i][triangle][corner][x, y]
    c = np.zeros((N1, N1, 2))
    c[:n, :n] = X[:, :, 0, 0]
    c[:n, n] = X[:, n - 1, 0, 1]
    c[n, :n] = X[n - 1, :, 1, 2]
    c[n, n] = X[n - 1, n - 1, 0, 2]
    # the reference joins periodic sides through the reflected partner point
    # (DGCell.cc:557-577, 687-693), which is the translation image only on a
    # domain centred at the origin: the mesh goes over shifted to be centred there
    # (a translation leaves every element's modal state and residual unchanged)
    c -= 0.5 * (c.reshape(-1, 2).min(axis=0) + c.reshape(-1, 2).max(axis=0))
    nid = lambda i, j: j * N1 + i + 1  # noqa: E731
    mp, sp = os.path.join(d, "m.msh"), os.path.join(d, "s.txt")
    with open(mp, "w") as f:
        f.write("$NOD\n%d\n" % (N1 * N1))
        for j in range(N1):
            for i in range(N1):
                f.write("%d %.17e %.17e 0\n" % (nid(i, j), c[j, i, 0], c[j, i, 1]))
        recs = []
        for j in range(n):
            recs += [(510, nid(0, j), nid(0, j + 1)), (510, nid(n, j), nid(n, j + 1))]
        for i in range(n):
            recs += [(610, nid(i, 0), nid(i + 1, 0)), (610, nid(i, n), nid(i + 1, n))]
        f.write("$ENDNOD\n$ELM\n%d\n" % (len(recs) + 2 * n * n))
        k = 1
        for tag, a, b in recs:
            f.write("%d 1 %d 1 2 %d %d\n" % (k, tag, a, b))
            k += 1
        tris = []
        for j in range(n):
            for i in range(n):
                tris.append((nid(i, j), nid(i + 1, j), nid(i + 1, j + 1)))
                tris.append((nid(i, j), nid(i + 1, j + 1), nid(i, j + 1)))
        for t in tris:
            f.write("%d 2 100 1 3 %d %d %d\n" % (k, t[0], t[1], t[2]))
            k += 1
        f.write("$ENDELM\n")
    Ue = np.asarray(U).reshape(len(tris), -1)
    with open(sp, "w") as f:
        for t, row in zip(tris, Ue):
            f.write("%d %d %d " % t + " ".join("%.17e" % x for x in row) + "\n")
    return mp, sp, len(tris)


def run_ref_driver(args, env):
    """oracle/_ref/ref_driver's stdout, or None when the build is absent or fails"""
    if not os.path.exists(REF_DRIVER):
        return None
    import subprocess
    r = subprocess.run([REF_DRIVER] + args, cwd=os.path.dirname(args[0]), env=dict(os.environ, **env),
                       capture_output=True, text=True, timeout=900)
    return r.stdout if r.returncode == 0 else None


def reference_cpu_baseline(mesh, U, dt, nsteps=1):
    """The reference's own serial code (oracle/_ref/ref_driver, built from the
    reference's sources with its Euler2d law and a local Lax-Friedrichs
    riemannSolver) on the full mesh: nsteps RK4 steps, timed by the driver
    around its step loop (mesh import and DGAnalysis construction excluded).
    None when the reference build is absent or fails."""
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        mp, sp, ntri = write_ref_lattice(mesh, NX, d, U)
        out = run_ref_driver([mp, "euler", "lf", str(ORDER), "rk4", sp, os.path.join(d, "o.txt"), "%r" % dt,
                              str(nsteps)], {"REF_FARFIELD": "1 1 1 3.5"})
    sec = [float(l.split()[1]) for l in (out or "").splitlines() if l.startswith("step_seconds ")]
    if not sec or sec[0] <= 0:
        return None
    return dict(value=np.asarray(U).size * 4 * nsteps / sec[0], unit="DOF*stage/s", cores=1, kind="reference",
                sample="%d RK4 step(s) of the full %d-triangle P3 vortex mesh by the reference's own serial "
                       "code (its DGAnalysis and TimeIntegrator built from its sources, LF flux), %.2f s of steps"
                % (nsteps, ntri, sec[0]))


TWO_PI = 2.0 * np.pi


def scalarpb_bump(x, y, z, t):
    """ScalarPb's initial condition (CaseOne, Lilia.cc:38-58): cos^2(2 pi r)
    within r <= 0.25 of (-0.5, 0)"""
    r = np.hypot(x + 0.5, y)
    return (np.cos(TWO_PI * r) ** 2 if r <= 0.25 else 0.0,)


def bytes_per_dof_stage_scalarpb():
    """Algorithmic HBM bytes per DOF and SSP-RK2 stage of the P1 scalar stage
    kernel with a velocity field (FP64): stage 1 reads U^n and writes U1, stage 2
    reads U1 and U^n and writes U^{n+1} (20 B per DOF on average); per element
    the table (|J| J^-1: 4 doubles, |J|, 3 int32 links) and the velocity at the
    3 volume and 3 x 2 face points (2 doubles each), over 3 DOF"""
    return (16 + 24) / 2.0 + (4 * 8 + 8 + 3 * 4) / 3.0 + (3 + 3 * 2) * 16 / 3.0


def run_scalarpb(dgpar, n=512, n_ref=128, steps=200, sweeps=40, tune=True, ref=True):
    """The reference's only benchmarked workload, ScalarPb (Lilia.cc:66-108):
    scalar advection with the rotating velocity FieldEvaluator (2 pi y, -2 pi x),
    CaseOne's bump as the initial state, P1 (SSP-RK2, the integrator the reference
    picks for order 1) at its default CFL 0.5 (DGAnalysis.cc:144), upwind flux,
    on jittered periodic triangle meshes of [-1,1]^2.  GPU: `steps` steps on
    n x n x 2 triangles (throughput), and the same on n_ref x n_ref x 2 (as many
    elements as the reference's advection16384 quad mesh, split); beside it the
    reference's own serial residual sweeps (volume + boundary contributions, as
    its profiling harness, DGAnalysis.cc:321-477) on that same mesh, timed on
    this host by oracle/_ref/ref_driver."""
    import tempfile
    out = {"workload": "ScalarPb: scalar advection, velocity FieldEvaluator (2 pi y, -2 pi x), CaseOne bump, P1, "
                       "upwind, SSP-RK2 CFL0.5, periodic [-1,1]^2 jittered triangles"}

    def velocity(x, y):
        return (TWO_PI * y, -TWO_PI * x)
    for key, m in (("gpu", n), ("gpu_reference_mesh", n_ref)):
        mesh = dgpar.mesh_tri_structured(m, m, x0=-1.0, y0=-1.0, lx=2.0, ly=2.0, jitter=0.1, seed=SEED)
        law = dgpar.Law(dgpar.LAW_SCALAR, dgpar.FLUX_UPWIND, farfield=(0.0,))
        g = dgpar.DGAnalysis(mesh, law, 1, precision=dgpar.PREC_FP64, use_graph=True, device=0)
        t0 = time.perf_counter()
        g.set_velocity_field(velocity)
        g.project_fn(scalarpb_bump)
        setup_s = time.perf_counter() - t0
        U0 = g.get_state()
        dt = g.computeDt(0.5)
        sched = tuned(g, dgpar.RK_TVD2, dt, tune)
        g.set_state(U0)
        g.time_advance(dgpar.RK_TVD2, dt, 10)
        g.set_state(U0)
        dgpar.device_synchronize()
        ms = g.time_advance(dgpar.RK_TVD2, dt, steps)
        total = float(g.norms()["totals"][0])
        g.set_state(U0)
        total0 = float(g.norms()["totals"][0])
        ndof = g.ndof
        g.close()
        bpd = bytes_per_dof_stage_scalarpb()
        ach = bpd * ndof / (ms / (2 * steps) / 1e3) / 1e9
        out[key] = {"mesh": "%dx%dx2 = %d triangles" % (m, m, 2 * m * m), "dof": ndof, "steps": steps, "dt": dt,
                    "ms_per_step": ms / steps, "value": ndof * 2 * steps / (ms / 1e3), "unit": "DOF*stage/s",
                    "bytes_per_dof_stage": bpd, "achieved_GBs": ach, "frac": ach / HBM_PEAK_GBS,
                    "field_tables_setup_s": setup_s, "schedule": sched,
                    # the periodic rotation conserves the total (int u)
                    "total_rel_change": abs(total - total0) / abs(total0)}
        if key == "gpu_reference_mesh" and ref:
            with tempfile.TemporaryDirectory() as d:
                mp, sp, ntri = write_ref_lattice(mesh, m, d, U0)
                so = run_ref_driver([mp, "scalar", "upwind", "1", "residual", sp, os.path.join(d, "o.txt")],
                                    {"REF_VEL": "rotate", "REF_FAR": "bump", "REF_SWEEPS": str(sweeps)})
            sec = [float(l.split()[1]) for l in (so or "").splitlines() if l.startswith("residual_seconds ")]
            if sec and sec[0] > 0:
                out["reference_cpu"] = {
                    "value": ndof * sweeps / sec[0], "unit": "DOF/s (residual sweeps: volume + boundary)",
                    "cores": 1, "kind": "reference", "sweeps": sweeps, "seconds": sec[0],
                    "sample": "%d serial residual sweeps of the reference's own code (ScalarLaw with the rotating "
                              "FieldEvaluator) on the same %d-triangle mesh" % (sweeps, ntri)}
                # a GPU stage is one residual plus the stage update per DOF
                out["gpu_stage_over_reference_sweep"] = out[key]["value"] / out["reference_cpu"]["value"]
    return out


def graded_mesh(dgpar, n, L, grade=(4, 2, 1, 1, 2, 4)):
    """n x n x 2 periodic triangles on [0, L]^2 over a tensor grid whose cell
    widths follow grade in blocks (the same in x and y): cells of several sizes
    for local time stepping"""
    w = np.repeat(np.asarray(grade, float), n // len(grade))
    w = np.concatenate([w, np.full(n - len(w), float(grade[-1]))])
    xs = np.concatenate([[0.0], np.cumsum(w)]) * (L / w.sum())
    vids = np.zeros((2 * n * n, 3), np.int32)
    xyz = np.zeros((2 * n * n, 3, 2))
    j, i = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    i, j = i.ravel(), j.ravel()
    topo = lambda a, b: (b % n) * (n + 1) + (a % n)  # noqa: E731
    for k, corners in enumerate((((0, 0), (1, 0), (1, 1)), ((0, 0), (1, 1), (0, 1)))):
        for v, (di, dj) in enumerate(corners):
            vids[k::2, v] = topo(i + di, j + dj)
            xyz[k::2, v, 0] = xs[i + di]
            xyz[k::2, v, 1] = xs[j + dj]
    return dgpar.Mesh(2, vids, xyz)


LTS_GRADE = (8,) * 10 + (4, 2, 1, 1, 2, 4) + (8,) * 10


def run_lts(dgpar, steps=40, n=256, levels=4):
    """Local time stepping (dg_lts_setup, DG_RK_TVD2_LTS): the isentropic vortex,
    P3, LF, on n x n x 2 periodic triangles of [0,10]^2 whose cell widths are
    8 outside a band through the centre graded 4:2:1:1:2:4, in both directions
    (a refined cross; four size levels).  LTS steps of the
    coarsest level's CFL 0.1 step h against global SSP-RK2 at the smallest
    element's CFL step over the same simulated time; the speedup is the ratio of
    wall time per simulated time, and the density L2 distance between the two
    end states is reported with the L2 error of each against the exact vortex."""
    mesh = graded_mesh(dgpar, n, 10.0, LTS_GRADE)
    law = dgpar.EulerLaw(dgpar.FLUX_LF, 1.4)
    out = {"workload": "euler2d-isentropic-vortex P3 LF, %dx%dx2 periodic [0,10]^2, cell widths 8 with a band graded "
                       "4:2:1:1:2:4 through the centre in x and y, local time stepping (at most %d size levels) vs "
                       "global SSP-RK2, CFL 0.1" % (n, n, levels)}
    with dgpar.DGAnalysis(mesh, law, ORDER, use_graph=True) as g:
        g.project(dgpar.FIELD_VORTEX, VORTEX)
        dt_min = g.computeDt(CFL)
        lev = g.sortCellsBySize(levels)
        h = g.computeDt(CFL)
        nlev = int(lev.max()) + 1
        g.time_advance(dgpar.RK_TVD2_LTS, h, 2)  # warm-up (graph capture)
        g.project(dgpar.FIELD_VORTEX, VORTEX)
        dgpar.device_synchronize()
        ms_lts = g.time_advance(dgpar.RK_TVD2_LTS, h, steps)
        T = steps * h
        err_lts = g.l2_error(dgpar.FIELD_VORTEX, VORTEX, T)
        U_lts = g.get_state()
        ndof = g.ndof
    nglob = int(np.ceil(T / dt_min - 1e-9))
    dtg = T / nglob
    with dgpar.DGAnalysis(mesh, law, ORDER, use_graph=True) as g:
        g.project(dgpar.FIELD_VORTEX, VORTEX)
        g.time_advance(dgpar.RK_TVD2, dtg, 2)
        g.project(dgpar.FIELD_VORTEX, VORTEX)
        dgpar.device_synchronize()
        ms_glob = g.time_advance(dgpar.RK_TVD2, dtg, nglob)
        err_glob = g.l2_error(dgpar.FIELD_VORTEX, VORTEX, T)
        U_glob = g.get_state()
    E = mesh.n_elems
    X = mesh.xyz
    detj = np.abs((X[:, 1, 0] - X[:, 0, 0]) * (X[:, 2, 1] - X[:, 0, 1]) -
                  (X[:, 2, 0] - X[:, 0, 0]) * (X[:, 1, 1] - X[:, 0, 1]))
    d = (U_lts - U_glob).reshape(E, 4, -1)[:, 0, :]
    counts = np.bincount(lev, minlength=nlev).tolist()
    # residual evaluations per element and step h: a level-b element needs at
    # least 2^(b+1) (its Heun substeps); global SSP-RK2 takes 2 h / dt per element
    out.update({"elements": E, "dof": ndof, "levels_histogram": counts, "h": h, "dt_min": dt_min,
                "lts_steps": steps, "lts_ms_per_step": ms_lts / steps, "simulated_time": T,
                "global_steps": nglob, "global_ms_per_step": ms_glob / nglob,
                "speedup_vs_global_ssp_rk2": ms_glob / ms_lts,
                "min_residuals_per_step": sum(c * 2 ** (b + 1) for b, c in enumerate(counts)) / E,
                "global_residuals_per_h": 2 * h / dtg,
                "l2_rho_vs_exact": {"lts": float(err_lts[0]), "global": float(err_glob[0])},
                "l2_rho_lts_vs_global": float(np.sqrt((d ** 2 * detj[:, None]).sum()))})
    return out


def run_hadapt(dgpar, steps=100, n=256, radius=2.0):
    """h-adaptivity (dg_adapt_h): the P3 LF isentropic vortex on the 256x256x2
    jittered periodic mesh of [0,10]^2; the cells within `radius` of the vortex
    centre are split twice (two passes, 1-irregular closure), which leaves
    hanging nodes on the patch's rim.  Reported: the wall time of each adapt
    pass (host refinement, projection and context rebuild), RK4 throughput on
    the adapted mesh (stage kernels + hanging-face passes) and the DOF share of
    each level; the totals' drift through the passes."""
    mesh = dgpar.mesh_tri_structured(n, n, lx=10.0, ly=10.0, jitter=0.1, seed=SEED)
    law = dgpar.EulerLaw(dgpar.FLUX_LF, 1.4)
    with dgpar.DGAnalysis(mesh, law, ORDER, use_graph=True) as g:
        g.project(dgpar.FIELD_VORTEX, VORTEX)
        t0 = g.norms()["totals"]
        adapt_s = []
        for _ in range(2):
            v, x, b, lev = g.refinedMesh()
            c = x.mean(axis=1)
            marks = (np.hypot(c[:, 0] - 5.0, c[:, 1] - 5.0) < radius).astype(np.int8)
            w0 = time.perf_counter()
            g.adaptH(marks, 2)
            adapt_s.append(time.perf_counter() - w0)
        v, x, b, lev = g.refinedMesh()
        drift = float(np.abs(g.norms()["totals"] - t0).max() / np.abs(t0).max())
        dt = g.computeDt(CFL)
        g.time_advance(dgpar.RK4, dt, 5)  # warm-up (graph capture)
        dgpar.device_synchronize()
        ms = g.time_advance(dgpar.RK4, dt, steps)
        ndof = g.ndof
    E = len(v)
    return {"workload": "euler2d-isentropic-vortex P3 LF RK4, 256x256x2 periodic [0,10]^2 jitter 0.1h, cells within "
                        "%.1f of the centre split twice (hanging nodes on the rim)" % radius,
            "elements": E, "dof": ndof, "levels_histogram": np.bincount(lev).tolist(),
            "adapt_pass_s": adapt_s, "totals_rel_drift": drift, "dt": dt, "steps": steps,
            "ms_per_step": ms / steps, "value": ndof * 4 * steps / (ms / 1e3), "unit": "DOF*stage/s"}


def cpu_model():
    """the host CPU's model name and the cores this process may use"""
    name = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    name = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"model": name, "cores_available": len(os.sched_getaffinity(0))}


def cpu_baseline(mesh, threads, budget_s=10.0):
    """RK4 steps of the FP64 CPU oracle on the same mesh on `threads` threads
    until about budget_s of CPU work (a bounded sample of the workload)"""
    from oracle import pyoracle as po
    o = po.Oracle(2, mesh.vids, mesh.xyz, ORDER, po.LAW_EULER2D, po.FLUX_LF, gamma=1.4)
    U = o.project(1, VORTEX)
    dt = o.dt(U, CFL)
    t0 = time.perf_counter()
    steps, el = 0, 0.0
    while el < budget_s:
        U, _ = o.step(U, po.RK4, steps * dt, dt, nthreads=threads)
        steps += 1
        el = time.perf_counter() - t0
    return dict(value=o.ndof * 4 * steps / el, unit="DOF*stage/s", cores=threads, kind="port",
                sample="%d RK4 steps of the full 131072-triangle P3 vortex mesh (FP64 CPU oracle, %d threads, %.1f s)"
                % (steps, threads, el))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--kernel", type=int, default=0, help="stage kernel: 0 tiled (default), 1 one element per lane")
    ap.add_argument("--adjoint-steps", type=int, default=200, help="RK4 steps of the adjoint leg (0: skip it)")
    ap.add_argument("--maxwell-steps", type=int, default=200,
                    help="RK4 steps of the Maxwell tetrahedra leg (0: skip it)")
    ap.add_argument("--mixed-steps", type=int, default=-1,
                    help="SSP-RK3 steps of the mixed-order limited shock-tube leg (0: skip it; -1: the whole run "
                         "to t = 0.2)")
    ap.add_argument("--scalar-steps", type=int, default=100,
                    help="SSP-RK3 steps of the scalar advection leg (0: skip it)")
    ap.add_argument("--scalarpb-steps", type=int, default=200,
                    help="SSP-RK2 steps of the ScalarPb leg (rotating velocity field; 0: skip it)")
    ap.add_argument("--lts-steps", type=int, default=40,
                    help="local-time-stepping steps of the graded-mesh leg (0: skip it)")
    ap.add_argument("--hadapt-steps", type=int, default=100,
                    help="RK4 steps on the h-adapted vortex mesh (0: skip it)")
    ap.add_argument("--no-tune", action="store_true", help="keep the default launch schedules (no tuner)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="first run a plain 256 MiB device copy (reference point for HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path; use --gpus 1", file=sys.stderr)
        return 2
    import dgpar

    if args.copy_probe:
        dgpar.copy_probe(64 << 20)
    mesh = dgpar.mesh_tri_structured(NX, NX, lx=10.0, ly=10.0, jitter=0.1, seed=SEED)
    E = mesh.n_elems
    tune = not args.no_tune
    r32 = run_form(dgpar, mesh, dgpar.PREC_FP32, args.steps, args.warmup, args.kernel, tune)
    r64 = run_form(dgpar, mesh, dgpar.PREC_FP64, args.steps, args.warmup, args.kernel, tune)
    ndof = r32["ndof"]
    np_ = (ORDER + 1) * (ORDER + 2) // 2
    X = mesh.xyz
    detj = np.abs((X[:, 1, 0] - X[:, 0, 0]) * (X[:, 2, 1] - X[:, 0, 1]) -
                  (X[:, 2, 0] - X[:, 0, 0]) * (X[:, 1, 1] - X[:, 0, 1]))
    # per-field L2 of FP32 - FP64: the orthonormal modal basis on the reference
    # triangle makes int_K u^2 = |J| sum_k c_k^2
    dd = (r32["state"] - r64["state"]).reshape(E, 4, np_)
    l2_diff = [float(np.sqrt((dd[:, f, :] ** 2 * detj[:, None]).sum())) for f in range(4)]
    l2_rho_diff = l2_diff[0]

    def thr(r):
        return ndof * 4 * args.steps / (r["ms"] / 1e3)

    v32, v64 = thr(r32), thr(r64)
    launch_ms = r32["ms"] / (4 * args.steps)
    b32 = bytes_per_dof_stage(True, 4, np_)
    b64 = bytes_per_dof_stage(False, 4, np_)
    ach = b32 * ndof / (launch_ms / 1e3) / 1e9
    ach64 = b64 * ndof / (r64["ms"] / (4 * args.steps) / 1e3) / 1e9
    out = {
        "metric": "DOF*stage/s of the FP32-arithmetic / FP64-accumulation DG residual + RK4 stage (vs FP64 plain)",
        "value": v32,
        "unit": "DOF*stage/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r32["ms"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 (fp64 stage accumulation)",
        "data": "synthetic: isentropic vortex projected on a seeded jittered mesh",
        "config": {"workload": "euler2d-isentropic-vortex P3 LF RK4 CFL0.1", "elements": E, "dof": ndof,
                   "mesh": "256x256x2 periodic [0,10]^2 jitter 0.1h seed %d" % SEED, "parallelism": "1 GPU"},
        "roofline": {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
                     "traffic": measured_traffic("fp32") if args.kernel == 0 else None,
                     "kernel": ("tri_stage_tiled<3,EulerLawD,float,*,%d>" % (8 if r32["schedule"]["best"][0] else 4)
                                if args.kernel == 0 else "tri_stage_kernel<3,EulerLawD,float,*>"),
                     "bytes_per_dof_stage": b32},
        "fp64_plain": {"value": v64, "ms_per_step": r64["ms"] / args.steps, "achieved_GBs": ach64,
                       "frac": ach64 / HBM_PEAK_GBS, "bytes_per_dof_stage": b64,
                       "traffic": measured_traffic("fp64") if args.kernel == 0 else None},
        "ratio_fp32_over_fp64": v32 / v64,
        "l2_rho_vs_exact": {"fp32": float(r32["err"][0]), "fp64": float(r64["err"][0])},
        "l2_rho_fp32_vs_fp64": l2_rho_diff,
        "l2_per_field": {"fields": ["rho", "rho u", "rho v", "rho E"],
                         "fp32_vs_exact": [float(x) for x in r32["err"]],
                         "fp64_vs_exact": [float(x) for x in r64["err"]],
                         "fp32_vs_fp64": l2_diff},
        "schedule": {"fp32": r32["schedule"], "fp64": r64["schedule"]},
        "wall_s_timed": r32["wall"],
    }
    if args.adjoint_steps > 0:
        out["adjoint"] = run_adjoint(dgpar, args.adjoint_steps, 20)
    if args.scalar_steps > 0:
        out["scalar_advection"] = run_scalar(dgpar, args.scalar_steps, tune=tune)
    if args.scalarpb_steps > 0:
        out["scalarpb"] = run_scalarpb(dgpar, steps=args.scalarpb_steps, tune=tune, ref=not args.no_cpu)
    if args.mixed_steps != 0:
        out["mixed_order_limited"] = run_mixed(dgpar, args.mixed_steps, tune=tune)
    if args.maxwell_steps > 0:
        out["maxwell3d_tets"] = run_maxwell(dgpar, args.maxwell_steps, tune=tune)
    if args.lts_steps > 0:
        out["local_time_stepping"] = run_lts(dgpar, args.lts_steps)
        # the same mesh with at most three levels: the finest cells join level
        # 2, so the step h halves, and the step holds 8 stage launches, not 16
        r3 = run_lts(dgpar, 2 * args.lts_steps, levels=3)
        out["local_time_stepping"]["three_levels"] = {
            k: r3[k] for k in ("levels_histogram", "h", "lts_steps", "lts_ms_per_step", "simulated_time",
                               "global_steps", "global_ms_per_step", "speedup_vs_global_ssp_rk2", "l2_rho_vs_exact")}
    if args.hadapt_steps > 0:
        out["h_adaptivity"] = run_hadapt(dgpar, args.hadapt_steps)
    if not args.no_cpu:
        # the reference's own serial code on the same mesh (its build travels with
        # the repository as oracle/_ref); the FP64 oracle port on this host's
        # cores beside it
        # the box's CPU share: 16 cores per GPU
        allc = max(1, min(16, len(os.sched_getaffinity(0))))
        port = {"threads_1": cpu_baseline(mesh, 1), "threads_%d" % allc: cpu_baseline(mesh, allc),
                "cpu": cpu_model()}
        from oracle import pyoracle as po
        o = po.Oracle(2, mesh.vids, mesh.xyz, ORDER, po.LAW_EULER2D, po.FLUX_LF, gamma=1.4)
        U0 = o.project(1, VORTEX)
        ref = reference_cpu_baseline(mesh, U0, o.dt(U0, CFL), nsteps=3)
        del o
        if ref is not None:
            ref["cpu"] = port["cpu"]
            out["cpu_baseline"] = ref
            out["cpu_baseline_port"] = port
        else:
            out["cpu_baseline"] = dict(port["threads_%d" % allc], cpu=port["cpu"])
            out["cpu_baseline_port"] = port
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
