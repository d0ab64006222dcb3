"""Local time stepping on the GPU (dg_lts_setup + DG_RK_TVD2_LTS steps) against
the oracle's restatement (oracle/dg_oracle.c or_lts_*), which evaluates every
stage of every element: the kernels skip the stages whose inputs repeat, so
agreement to rounding checks the skipping exactly.  The oracle is pinned in
tests/test_lts.py (SSP-RK2 on one level, the partitioned Runge-Kutta tableaux,
conservation, second order)."""
import numpy as np
import pytest

import dgpar
from helpers import graded_nodes, mesh_tri, mesh_tri_graded
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

VORTEX = (5.0, 5.0, 5.0, 1.0, 1.0, 10.0)


def graded(n, lo=0.0, hi=1.0, grade=(4, 2, 1, 1, 2, 4)):
    xs = graded_nodes(n, lo, hi, grade)
    return mesh_tri_graded(xs, xs)


def case(law, p, n=12, flux=None):
    if law == "euler":
        vids, xyz = graded(n, 0.0, 10.0)
        fl = flux if flux is not None else dgpar.FLUX_LF
        o = po.Oracle(2, vids, xyz, p, po.LAW_EULER2D, {dgpar.FLUX_LF: po.FLUX_LF, dgpar.FLUX_ROE: po.FLUX_ROE}[fl],
                      gamma=1.4)
        lw = dgpar.EulerLaw(fl, 1.4)
        U = o.project(1, VORTEX)
    else:
        vids, xyz = graded(n)
        o = po.Oracle(2, vids, xyz, p, po.LAW_SCALAR, po.FLUX_UPWIND, vel=(1.0, 0.5))
        lw = dgpar.Law(dgpar.LAW_SCALAR, dgpar.FLUX_UPWIND, vel=(1.0, 0.5), farfield=(0.0,))
        U = o.project(0)
    return dgpar.Mesh(2, vids, xyz), lw, o, U


@pytest.mark.parametrize("law,p,ml,shape", [("scalar", 2, 3, 0), ("scalar", 1, 4, 0), ("euler", 1, 3, 0),
                                             ("euler", 3, 3, 0), ("euler", 2, 2, 1), ("euler", 4, 3, 0)])
def test_lts_steps_match_oracle(law, p, ml, shape):
    print("lts", law, p, ml, shape, flush=True)
    mesh, lw, o, U = case(law, p)
    lev_o, n = o.lts_levels(ml)
    with dgpar.DGAnalysis(mesh, lw, p, use_graph=False) as g:
        g.set_state(U)
        lev = g.sortCellsBySize(ml)
        np.testing.assert_array_equal(lev, lev_o)
        # the state survives the re-layout
        np.testing.assert_array_equal(g.get_state(), U)
        g.set_schedule(shape, 0)
        h = g.computeDt(0.1)
        ho = o.lts_dt(U, 0.1, lev_o)
        assert abs(h - ho) <= 1e-14 * ho
        g.advanceInTime(0.0, h, dgpar.RK_TVD2_LTS, 3)
        V = U
        for _ in range(3):
            V = o.lts_step(V, lev_o, h)
        Ug = g.get_state()
    assert np.abs(Ug - V).max() <= 1e-12 * np.abs(V).max()


def test_one_level_equals_ssp_rk2():
    vids, xyz, _ = mesh_tri(8, 8, jitter=0.1, seed=4)
    mesh = dgpar.Mesh(2, vids, xyz)
    lw = dgpar.Law(dgpar.LAW_SCALAR, dgpar.FLUX_UPWIND, vel=(1.0, 0.5), farfield=(0.0,))
    with dgpar.DGAnalysis(mesh, lw, 2) as a, dgpar.DGAnalysis(mesh, lw, 2) as b:
        a.project(dgpar.FIELD_GAUSSIAN)
        b.project(dgpar.FIELD_GAUSSIAN)
        assert not b.sortCellsBySize(1).any()
        h = a.computeDt(0.1)
        assert b.computeDt(0.1) == h
        a.advanceInTime(0.0, h, dgpar.RK_TVD2, 5)
        b.advanceInTime(0.0, h, dgpar.RK_TVD2_LTS, 5)
        Ua, Ub = a.get_state(), b.get_state()
    assert np.abs(Ua - Ub).max() <= 1e-14 * np.abs(Ua).max()


def test_graph_equals_eager_and_conserves():
    mesh, lw, o, U = case("euler", 2, flux=dgpar.FLUX_ROE)
    res = []
    for graph in (False, True):
        with dgpar.DGAnalysis(mesh, lw, 2, use_graph=graph) as g:
            g.set_state(U)
            g.sortCellsBySize(3)
            h = g.computeDt(0.1)
            t0 = g.norms()["totals"]
            for _ in range(4):  # graph replays with host work between them
                g.advanceInTime(0.0, h, dgpar.RK_TVD2_LTS, 5)
            res.append(g.get_state())
            np.testing.assert_allclose(g.norms()["totals"], t0, rtol=1e-13, atol=1e-13)
    np.testing.assert_array_equal(res[0], res[1])


def test_graph_several_steps_and_new_plan():
    """70 steps run eight eight-step graphs and six one-step graphs; a new level
    plan on the same context, stepped with the same dt, drops the captured steps
    of the old plan.  Both agree with eager steps bit for bit."""
    mesh, lw, o, U = case("euler", 2, flux=dgpar.FLUX_ROE)
    res = []
    for graph in (False, True):
        with dgpar.DGAnalysis(mesh, lw, 2, use_graph=graph) as g:
            g.set_state(U)
            g.sortCellsBySize(2)
            h2 = g.computeDt(0.1)
            g.sortCellsBySize(3)
            h = min(h2, g.computeDt(0.1))
            g.advanceInTime(0.0, h, dgpar.RK_TVD2_LTS, 70)
            g.sortCellsBySize(2)
            g.advanceInTime(0.0, h, dgpar.RK_TVD2_LTS, 70)
            res.append(g.get_state())
    np.testing.assert_array_equal(res[0], res[1])


def test_refusals():
    mesh, lw, o, U = case("scalar", 1)
    with dgpar.DGAnalysis(mesh, lw, 1) as g:
        with pytest.raises(dgpar.DGError):
            g.advanceInTime(0.0, 1e-3, dgpar.RK_TVD2_LTS, 1)  # no setup
        g.sortCellsBySize(3)
        g.set_limiter(True, 0.0, 1.0)
        with pytest.raises(dgpar.DGError):
            g.advanceInTime(0.0, 1e-3, dgpar.RK_TVD2_LTS, 1)
    with dgpar.DGAnalysis(mesh, lw, 1, precision=dgpar.PREC_FP32) as g:
        with pytest.raises(dgpar.DGError):
            g.sortCellsBySize(3)


def test_large_graded_mesh_one_step():
    # 96 x 96 x 2 = 18 432 triangles, cell widths graded 4:1 in both directions
    mesh, lw, o, U = case("euler", 3, n=96)
    lev_o, n = o.lts_levels(3)
    assert n == 3
    with dgpar.DGAnalysis(mesh, lw, 3) as g:
        g.set_state(U)
        g.sortCellsBySize(3)
        h = g.computeDt(0.1)
        g.advanceInTime(0.0, h, dgpar.RK_TVD2_LTS, 1)
        Ug = g.get_state()
    V = o.lts_step(U, lev_o, h, nthreads=8)
    assert np.abs(Ug - V).max() <= 1e-12 * np.abs(V).max()
