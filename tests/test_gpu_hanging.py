"""Meshes with hanging nodes and h-adaptivity on the GPU: residuals and RK steps
against the oracle (whose faces of half an edge are pinned in
tests/test_hanging.py), dg_adapt_h's meshes against an independent red
refinement (tests/helpers.py refine_tri), exact state transfer and
conservation, the 1-irregular closure, and the refusals."""
import numpy as np
import pytest

import dgpar
from helpers import mesh_tri, refine_tri
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

VORTEX = (5.0, 5.0, 5.0, 1.0, 1.0, 10.0)
FF = (1.0, 0.1, 0.05, 2.5)


def split_mesh(n=8, periodic=True, pick=(9, 10, 27, 40, 41, 60, 100), lx=10.0, seed=7):
    vids, xyz, bc = mesh_tri(n, n, lx=lx, ly=lx, periodic_x=periodic, periodic_y=periodic, jitter=0.1, seed=seed,
                             side_tags=(7, 7, 40000, 7))
    mask = np.zeros(len(vids), bool)
    mask[[p for p in pick if p < len(vids)]] = True
    return refine_tri(vids, xyz, bc, mask)


def laws(law):
    if law == "scalar":
        return (dgpar.Law(dgpar.LAW_SCALAR, dgpar.FLUX_UPWIND, vel=(1.0, 0.5), farfield=(0.2,)),
                dict(law=po.LAW_SCALAR, flux=po.FLUX_UPWIND, vel=(1.0, 0.5), farfield=(0.2,)))
    if law == "maxwell2d":
        return (dgpar.Maxwell2dLaw(farfield=(0.1, 0.0, 0.2)),
                dict(law=po.LAW_MAXWELL2D, flux=po.FLUX_UPWIND, farfield=(0.1, 0.0, 0.2)))
    if law == "burgers":
        return (dgpar.Burgers2DLaw((0.4,)), dict(law=po.LAW_BURGERS2D, flux=po.FLUX_CENTRAL, farfield=(0.4,)))
    fl = {"lf": (dgpar.FLUX_LF, po.FLUX_LF), "roe": (dgpar.FLUX_ROE, po.FLUX_ROE),
          "exact": (dgpar.FLUX_EXACT, po.FLUX_EXACT)}[law]
    return (dgpar.Law(dgpar.LAW_EULER2D, fl[0], gamma=1.4, farfield=FF),
            dict(law=po.LAW_EULER2D, flux=fl[1], gamma=1.4, farfield=FF))


def oracle(v, x, b, p, od):
    return po.Oracle(2, v, x, p, od["law"], od["flux"], gamma=od.get("gamma", 1.4), vel=od.get("vel", (0, 0, 0)),
                     farfield=od.get("farfield", ()), bc=b)


def state(o, law, seed=3):
    if law == "scalar":
        U = o.project_fn(lambda X, Y, Z, t: (np.exp(-((X - 5) ** 2 + (Y - 5) ** 2) / 4.0),))
    elif law == "maxwell2d":
        U = o.project_fn(lambda X, Y, Z, t: (np.sin(0.6 * Y), np.cos(0.5 * X), np.sin(0.4 * (X + Y))))
    elif law == "burgers":
        U = o.project_fn(lambda X, Y, Z, t: (0.5 + 0.3 * np.sin(0.6 * X) * np.cos(0.4 * Y),))
    else:
        U = o.project(1, VORTEX)
    return U + 0.002 * np.random.default_rng(seed).standard_normal(U.shape)


@pytest.mark.parametrize("law,p,periodic", [("scalar", 1, True), ("scalar", 3, False), ("lf", 1, False),
                                             ("lf", 3, True), ("roe", 2, True), ("roe", 4, False),
                                             ("exact", 2, True), ("maxwell2d", 3, False), ("burgers", 2, True),
                                             ("burgers", 3, False)])
def test_residual_matches_oracle(law, p, periodic):
    print("hanging residual", law, p, periodic, flush=True)
    v, x, b, _ = split_mesh(periodic=periodic)
    lw, od = laws(law)
    o = oracle(v, x, b, p, od)
    U = state(o, law)
    with dgpar.DGAnalysis(dgpar.Mesh(2, v, x, b), lw, p) as g:
        g.set_state(U)
        R = g.residual()
        F = g.faces()
    Ro, _ = o.residual(U)
    assert np.abs(R - Ro).max() <= 1e-12 * np.abs(Ro).max()
    # the faces: the oracle's set, halves with the child on the left
    key = lambda f: tuple(f)  # noqa: E731
    assert sorted(map(key, F)) == sorted(map(key, o.faces()))


@pytest.mark.parametrize("scheme,shape", [(dgpar.RK4, 0), (dgpar.RK_TVD3, 1), (dgpar.RK_TVD2, 0), (dgpar.RK_FE, 0)])
def test_steps_match_oracle(scheme, shape):
    v, x, b, _ = split_mesh()
    lw, od = laws("lf")
    o = oracle(v, x, b, 3, od)
    U = state(o, "lf")
    dt = o.dt(U, 0.1)
    for graph in (False, True):
        with dgpar.DGAnalysis(dgpar.Mesh(2, v, x, b), lw, 3, use_graph=graph) as g:
            g.set_schedule(shape, 0)
            g.set_state(U)
            g.advanceInTime(0.0, dt, scheme, 3)
            W = g.get_state()
        V = U
        for k in range(3):
            V, _ = o.step(V, scheme, k * dt, dt)
        assert np.abs(W - V).max() <= 1e-12 * np.abs(V).max()


def test_adapt_split_matches_independent_refinement_and_roundtrips():
    vids, xyz, bc = mesh_tri(8, 8, lx=10.0, ly=10.0, jitter=0.1, seed=7)
    lw, od = laws("lf")
    with dgpar.DGAnalysis(dgpar.Mesh(2, vids, xyz, bc), lw, 3) as g:
        g.project(dgpar.FIELD_VORTEX, VORTEX)
        U0 = g.get_state()
        t0 = g.norms()["totals"]
        marks = np.zeros(len(vids), np.int8)
        marks[[9, 10, 27, 40, 41, 60, 100]] = 1
        n = g.adaptH(marks, 1)
        v2, x2, b2, lev = g.refinedMesh()
        rv, rx, rb, par = refine_tri(vids, xyz, bc, marks > 0)
        assert n == len(rv)
        np.testing.assert_array_equal(v2, rv)
        np.testing.assert_allclose(x2, rx, rtol=0, atol=1e-15)
        np.testing.assert_array_equal(b2, rb)
        np.testing.assert_array_equal(lev, (marks > 0)[par].astype(np.int32))
        # the projection keeps each cell's polynomial: totals, and unsplit cells' states
        np.testing.assert_allclose(g.norms()["totals"], t0, rtol=1e-13)
        U1 = g.get_state().reshape(n, -1)
        keep = ~(marks > 0)[par]
        np.testing.assert_array_equal(U1[keep], U0.reshape(len(vids), -1)[par[keep]])
        # residual parity on the adapted mesh
        o = oracle(v2, x2, b2, 3, od)
        Ro, _ = o.residual(g.get_state())
        assert np.abs(g.residual() - Ro).max() <= 1e-12 * np.abs(Ro).max()
        # merging every family back restores the mesh and the state
        back = np.where(lev > 0, -1, 0).astype(np.int8)
        assert g.adaptH(back, 1) == len(vids)
        v3, x3, b3, lev3 = g.refinedMesh()
        np.testing.assert_array_equal(v3, vids)
        assert not lev3.any()
        assert np.abs(g.get_state() - U0).max() <= 1e-13 * np.abs(U0).max()


def test_adapt_two_levels_closure_and_merge_rules():
    vids, xyz, bc = mesh_tri(8, 8, lx=10.0, ly=10.0, jitter=0.1, seed=7)
    lw, od = laws("roe")
    with dgpar.DGAnalysis(dgpar.Mesh(2, vids, xyz, bc), lw, 2) as g:
        g.project(dgpar.FIELD_VORTEX, VORTEX)
        t0 = g.norms()["totals"]
        m = np.zeros(len(vids), np.int8)
        m[[40, 41]] = 1
        g.adaptH(m, 2)
        v, x, b, lev = g.refinedMesh()
        # split a child on the edge of the refined patch: its coarse neighbours
        # must be split too (1-irregular closure)
        F = g.faces()
        kid = next(e for e in range(len(v)) if lev[e] == 1 and any(lev[o] == 0 for o in
                   np.concatenate([F[F[:, 0] == e, 2], F[F[:, 2] == e, 0]]) if o >= 0))
        m2 = np.zeros(len(v), np.int8)
        m2[kid] = 1
        g.adaptH(m2, 2)
        v, x, b, lev = g.refinedMesh()
        assert lev.max() == 2
        F = g.faces()
        inner = F[F[:, 2] >= 0]
        assert np.abs(lev[inner[:, 0]] - lev[inner[:, 2]]).max() <= 1
        np.testing.assert_allclose(g.norms()["totals"], t0, rtol=1e-13)
        o = oracle(v, x, b, 2, od)
        U = g.get_state()
        Ro, _ = o.residual(U)
        assert np.abs(g.residual() - Ro).max() <= 1e-12 * np.abs(Ro).max()
        dt = g.computeDt(0.1)
        g.advanceInTime(0.0, dt, dgpar.RK_TVD3, 2)
        V = U
        for k in range(2):
            V, _ = o.step(V, po.RK_TVD3, k * dt, dt)
        assert np.abs(g.get_state() - V).max() <= 1e-12 * np.abs(V).max()
        # coarsening everything: a family next to a cell two levels finer waits,
        # so the first pass merges the level-2 families and the level-1 families
        # away from them, the second the rest, which restores the first mesh
        for k in range(2):
            v4, _, _, lev4 = g.refinedMesh()
            g.adaptH(np.where(lev4 > 0, -1, 0).astype(np.int8), 2)
            v5, _, _, lev5 = g.refinedMesh()
            F5 = g.faces()
            inner = F5[F5[:, 2] >= 0]
            assert np.abs(lev5[inner[:, 0]] - lev5[inner[:, 2]]).max() <= 1
            np.testing.assert_allclose(g.norms()["totals"], t0, rtol=1e-13)
            if k == 0:
                assert lev5.max() == 1 and len(v5) < len(v4)
        assert not lev5.any()
        np.testing.assert_array_equal(v5, vids)


def test_refusals_on_hanging_meshes():
    v, x, b, _ = split_mesh()
    lw, od = laws("lf")
    with dgpar.DGAnalysis(dgpar.Mesh(2, v, x, b), lw, 2) as g:
        for call in (lambda: g.set_limiter(True, 0.0, 1.0), lambda: g.set_recovery(True),
                     lambda: g.sortCellsBySize(3), lambda: g.residual_jvp(g.get_state(), g.get_state())):
            with pytest.raises(dgpar.DGError):
                call()
    with pytest.raises(dgpar.DGError):
        dgpar.DGAnalysis(dgpar.Mesh(2, v, x, b), lw, 2, precision=dgpar.PREC_FP32)


def test_large_adapted_mesh():
    # 128 x 128 x 2 periodic vortex mesh, the cells within 2 of the vortex centre
    # split twice: residual and one RK4 step against the oracle, totals kept
    vids, xyz, bc = mesh_tri(128, 128, lx=10.0, ly=10.0, jitter=0.1, seed=11)
    lw, od = laws("lf")
    with dgpar.DGAnalysis(dgpar.Mesh(2, vids, xyz, bc), lw, 3) as g:
        g.project(dgpar.FIELD_VORTEX, VORTEX)
        t0 = g.norms()["totals"]
        for _ in range(2):
            v, x, b, lev = g.refinedMesh()
            c = x.mean(axis=1)
            marks = (np.hypot(c[:, 0] - 5.0, c[:, 1] - 5.0) < 2.0).astype(np.int8)
            g.adaptH(marks, 2)
        v, x, b, lev = g.refinedMesh()
        assert lev.max() == 2 and len(v) > 40000
        np.testing.assert_allclose(g.norms()["totals"], t0, rtol=1e-12)
        U = g.get_state()
        o = oracle(v, x, b, 3, od)
        Ro, _ = o.residual(U, nthreads=8)
        # small cells on [0, 10]^2: the residual (~|J|) is far below its terms'
        # scale, so the bound is relative to the summed term magnitudes
        scale = o.residual_scale(U, nthreads=8).max()
        err = np.abs(g.residual() - Ro).max()
        print("large adapted mesh: |dR| %.3e  max|R| %.3e  term scale %.3e" % (err, np.abs(Ro).max(), scale))
        assert err <= 1e-12 * scale
        dt = g.computeDt(0.1)
        g.advanceInTime(0.0, dt, dgpar.RK4, 1)
        V, _ = o.step(U, po.RK4, 0.0, dt, nthreads=8)
        assert np.abs(g.get_state() - V).max() <= 1e-12 * np.abs(V).max()
