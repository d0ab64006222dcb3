"""GPU: non-physical states through the C ABI (dg_recover, dg_set_recovery and
dg_advance's check after its steps), against the oracle's recovery
(DGCell::setToZero semantics):

- flags bit for bit and the recovered state equal to the oracle's on the same
  state (uniform P3 and mixed P1-P4);
- dg_advance reports DG_ERR_NONPHYSICAL for a state left non-physical and
  DG_ERR_NONFINITE for a NaN, instead of stepping on silently;
- with recovery on, limited or not, RK steps of a state with non-physical
  elements match the oracle stepping with recovery (1e-12) and succeed."""
import sys

import numpy as np
import pytest

from helpers import mesh_tri, np_tri
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

FF = (1.0, 0.2, -0.1, 2.6)


def announce(name):
    print("\n[gpu case] %s" % name, flush=True)
    sys.stdout.flush()


def make(orders_seed=None, p=3, n=10, use_graph=True):
    import dgpar
    vids, xyz, bc = mesh_tri(n, n, lx=1.0, ly=1.0, periodic_x=False, periodic_y=False, jitter=0.1, seed=4,
                             side_tags=(40000, 1, 40000, 1))
    E = vids.shape[0]
    orders = (np.random.default_rng(orders_seed).integers(1, 5, E).astype(np.int32) if orders_seed is not None
              else np.full(E, p, np.int32))
    o = po.Oracle(2, vids, xyz, orders, po.LAW_EULER2D, po.FLUX_ROE_EF, gamma=1.4, farfield=FF, bc=bc)
    g = dgpar.DGAnalysis(dgpar.Mesh(2, vids, xyz, bc, order=None if orders_seed is None else orders),
                         dgpar.Law(dgpar.LAW_EULER2D, dgpar.FLUX_ROE_EF, gamma=1.4, farfield=FF), p,
                         use_graph=use_graph)
    return o, g, orders


def state(orders, seed, nbad=8):
    """free stream plus 1% noise above the means; nbad elements pushed far below
    zero density at some points through their modes above the mean"""
    rng = np.random.default_rng(seed)
    bad = set(rng.choice(len(orders), nbad, replace=False).tolist())
    out = []
    for e, p in enumerate(orders):
        b = np.zeros((4, np_tri(p)))
        b[:, 0] = np.array(FF) / np.sqrt(2.0)
        b[:, 1:] = 0.01 * rng.standard_normal((4, np_tri(p) - 1))
        if e in bad:
            b[0, 1:] = 1.5 * rng.standard_normal(np_tri(p) - 1)
        out.append(b.ravel())
    return np.concatenate(out), sorted(bad)


@pytest.mark.parametrize("seed", [None, 3])
def test_recover_flags_and_state_match_the_oracle(seed):
    announce("recover flags/state orders seed %s" % seed)
    o, g, orders = make(seed)
    U, bad = state(orders, 7)
    g.set_state(U)
    f_check = g.recover(apply=False)
    assert np.array_equal(g.get_state(), U)  # a check changes nothing
    Uo, fo = o.recover(U)
    assert np.array_equal(f_check, fo) and (fo & 1).sum() >= 4
    assert set(np.nonzero(fo)[0]) <= set(bad)
    f = g.recover(apply=True)
    assert np.array_equal(f, fo)
    assert np.array_equal(g.get_state(), Uo)
    assert np.array_equal(g.recovery_flags(), fo)
    assert not g.recover(apply=False).any()


def test_recover_reports_a_nonphysical_mean():
    announce("recover: non-physical mean")
    o, g, orders = make(None, p=2)
    U = np.zeros((len(orders), 4, 6))
    U[:, :, 0] = np.array(FF) / np.sqrt(2.0)
    U[3, 0, 0] = -0.1
    U[3, 0, 1] = 0.5
    U = U.ravel()
    g.set_state(U)
    f = g.recover(apply=False)
    _, fo = o.recover(U)
    assert np.array_equal(f, fo) and f[3] == 3 and np.count_nonzero(f) == 1


def test_advance_reports_nonphysical_and_nonfinite_states():
    import dgpar
    announce("advance: NONPHYSICAL and NONFINITE")
    o, g, orders = make(None, p=3)
    U, bad = state(orders, 11)
    g.set_state(U)
    with pytest.raises(dgpar.DGError) as ei:
        g.advanceInTime(0.0, 1e-9, dgpar.RK_TVD3, 1)
    assert ei.value.status == dgpar.DG_ERR_NONPHYSICAL
    # dg_set_state refuses NaN; an overflowing step makes them: a physical state
    # advanced with dt = 1e300
    W = o.recover(U)[0]
    g.set_state(W)
    g.advanceInTime(0.0, 1e-4, dgpar.RK_TVD3, 2)  # a physical state steps without complaint
    g.set_state(W)
    with pytest.raises(dgpar.DGError) as ei:
        g.advanceInTime(0.0, 1e300, dgpar.RK_TVD3, 1)
    assert ei.value.status == dgpar.DG_ERR_NONFINITE
    X = g.get_state()
    assert not np.isfinite(X).all()
    f = g.recover(apply=False)
    _, fo = o.recover(X)
    assert np.array_equal(f, fo) and (f & 4).any()


@pytest.mark.parametrize("seed,limiter,use_graph", [(None, False, True), (None, True, False), (5, False, True),
                                                    (5, True, True)])
def test_recovered_steps_match_the_oracle(seed, limiter, use_graph):
    import dgpar
    announce("recovered steps seed %s limiter %d graph %d" % (seed, limiter, use_graph))
    o, g, orders = make(seed, use_graph=use_graph)
    U, bad = state(orders, 13)
    dt = 0.5 * o.dt(o.recover(U)[0], 0.1)
    o.set_recovery(True)
    o.set_limiter(limiter)
    Uo, t = U.copy(), 0.0
    for _ in range(4):
        Uo, ro = o.step(Uo, po.RK_TVD3, t, dt)
        t += dt
    g.set_state(U)
    g.set_recovery(True)
    g.set_limiter(limiter)
    r = g.advanceInTime(0.0, dt, dgpar.RK_TVD3, 4)
    assert np.abs(g.get_state() - Uo).max() <= 1e-12 * np.abs(Uo).max()
    assert abs(r - ro) <= 1e-10 * abs(ro)
    assert g.recovery_flags().any()


def test_recovery_refuses_what_it_does_not_serve():
    import dgpar
    announce("recovery argument errors")
    m = dgpar.mesh_tri_structured(4, 4)
    g = dgpar.DGAnalysis(m, dgpar.Law(dgpar.LAW_SCALAR, dgpar.FLUX_UPWIND, vel=(1, 0)), 2)
    with pytest.raises(dgpar.DGError):
        g.set_recovery(True)
    h = dgpar.DGAnalysis(m, dgpar.EulerLaw(dgpar.FLUX_LF, 1.4), 2, precision=dgpar.PREC_FP32)
    with pytest.raises(dgpar.DGError):
        h.set_recovery(True)
    t = dgpar.DGAnalysis(dgpar.mesh_tet_structured(2, 2, 2), dgpar.MaxwellLaw(), 1)
    with pytest.raises(dgpar.DGError):
        t.set_recovery(True)
    with pytest.raises(dgpar.DGError):
        t.recover(apply=False)
