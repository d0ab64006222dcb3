"""GPU: launch schedules and the tuner through the C ABI.

- the 8-wave tile shape of the triangle kernel (P1-P4, scalar and Euler) and
  of the tetrahedron kernel (P1-P3) against the oracle within 1e-12;
- a schedule changes no state: RK steps under every candidate schedule are bit
  for bit those of the default schedule (triangles, tetrahedra, the persistent
  mixed-order launch; its bucket-by-bucket form, compiled apart, to FP64
  rounding); the step residual norm, a sum over workgroups and their
  threads, agrees within 1e-14 (its summation order follows the launch);
- the tuner restores the state it timed on, picks one of its candidates, and a
  second context of the same configuration reads the choice back from the cache
  file without timing;
- schedules a context cannot run are refused."""
import sys

import numpy as np
import pytest

from helpers import mesh_tri, np_tri
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

def announce(name):
    print("\n[gpu case] %s" % name, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def tri_pair(p, law="euler", n=8, flux=po.FLUX_ROE, precision=0):
    import dgpar
    vids, xyz, bc = mesh_tri(n, n, lx=1.0, ly=1.0, periodic_x=False, periodic_y=True, jitter=0.1, seed=3,
                             side_tags=(7, 7, 7, 7))
    if law == "euler":
        ff = (1.0, 0.3, 0.1, 2.6)
        o = po.Oracle(2, vids, xyz, p, po.LAW_EULER2D, flux, gamma=1.4, farfield=ff, bc=bc)
        lw = dgpar.Law(dgpar.LAW_EULER2D, flux, gamma=1.4, farfield=ff)
    else:
        o = po.Oracle(2, vids, xyz, p, po.LAW_SCALAR, po.FLUX_UPWIND, vel=(1.0, 0.5), farfield=(0.2,), bc=bc)
        lw = dgpar.Law(dgpar.LAW_SCALAR, dgpar.FLUX_UPWIND, vel=(1.0, 0.5), farfield=(0.2,))
    g = dgpar.DGAnalysis(dgpar.Mesh(2, vids, xyz, bc), lw, p, precision=precision)
    return o, g


def smooth_euler(o, p, E, seed):
    rng = np.random.default_rng(seed)
    b = 0.01 * rng.standard_normal((E, 4, np_tri(p)))
    b[:, :, 0] += np.array([1.0, 0.2, -0.1, 2.5]) / np.sqrt(2.0)  # means, phi_0 = sqrt(2)
    return b.ravel()


@pytest.mark.parametrize("law", ["scalar", "euler"])
@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_tri_8_wave_tiles_match_oracle(p, law):
    announce("tri 8-wave residual P%d %s" % (p, law))
    o, g = tri_pair(p, law)
    U = (smooth_euler(o, p, o.ndof // (4 * np_tri(p)), p) if law == "euler"
         else np.random.default_rng(p).standard_normal(o.ndof))
    g.set_state(U)
    g.set_schedule(shape=1)
    assert g.schedule() == (1, 0)
    assert rel(g.residual(), o.residual(U)[0]) < 1e-12


@pytest.mark.parametrize("p", [1, 2, 3])
def test_tet_8_wave_tiles_match_oracle(p):
    import dgpar
    announce("tet 8-wave residual P%d" % p)
    m = dgpar.mesh_tet_structured(3, 3, 3)
    o = po.Oracle(3, m.vids, m.xyz, p, po.LAW_MAXWELL3D, po.FLUX_UPWIND)
    g = dgpar.DGAnalysis(m, dgpar.MaxwellLaw(po.FLUX_UPWIND), p)
    U = np.random.default_rng(30 + p).standard_normal(o.ndof)
    g.set_state(U)
    g.set_schedule(shape=1)
    assert rel(g.residual(), o.residual(U)[0]) < 1e-12


def steps_under(g, U, dt, scheme, n, shape=0, grid=0):
    g.set_state(U)
    g.set_schedule(shape=shape, grid=grid)
    r = g.advanceInTime(0.0, dt, scheme=scheme, nsteps=n)
    return g.get_state(), r


@pytest.mark.parametrize("precision", [0, 1])
def test_tri_schedules_give_identical_steps(precision):
    announce("tri schedules identical, precision %d" % precision)
    p = 3
    o, g = tri_pair(p, precision=precision)
    U = smooth_euler(o, p, o.ndof // (4 * np_tri(p)), 5)
    dt = o.dt(U, 0.1)
    U0, r0 = steps_under(g, U, dt, po.RK4, 4, shape=0)
    U1, r1 = steps_under(g, U, dt, po.RK4, 4, shape=1)
    assert np.array_equal(U0, U1) and abs(r0 - r1) <= 1e-14 * abs(r0)


def test_tet_schedules_give_identical_steps():
    import dgpar
    announce("tet schedules identical")
    m = dgpar.mesh_tet_structured(4, 4, 4)
    g = dgpar.DGAnalysis(m, dgpar.MaxwellLaw(po.FLUX_UPWIND), 3)
    g.project(3)
    U = g.get_state()
    dt = g.computeDt(0.1)
    U0, r0 = steps_under(g, U, dt, po.RK4, 3, shape=0)
    U1, r1 = steps_under(g, U, dt, po.RK4, 3, shape=1)
    assert np.array_equal(U0, U1) and abs(r0 - r1) <= 1e-14 * abs(r0)


def mixed_ctx(nx=64, ny=32, seed=21):
    import dgpar
    sod = (1.0, 0.0, 0.0, 2.5, 0.125, 0.0, 0.0, 0.25)
    vids, xyz, bc = mesh_tri(nx, ny, lx=1.0, ly=0.25, periodic_x=False, periodic_y=False, jitter=0.1, seed=4,
                             side_tags=(40000, 2, 40000, 1))
    orders = np.random.default_rng(seed).integers(1, 5, vids.shape[0]).astype(np.int32)
    g = dgpar.DGAnalysis(dgpar.Mesh(2, vids, xyz, bc, order=orders),
                         dgpar.Law(dgpar.LAW_EULER2D, dgpar.FLUX_ROE_EF, gamma=1.4, farfield=sod), 1)
    return g


def test_mixed_grids_give_identical_steps():
    announce("mixed grids identical")
    g = mixed_ctx()
    g.project(2, (0.53,))
    g.set_limiter(True)
    U = g.get_state()
    dt = 0.5 * g.computeDt(0.1)
    ref, r0 = steps_under(g, U, dt, po.RK_TVD3, 3)
    for grid in (1, 3, 7):
        Ug, rg = steps_under(g, U, dt, po.RK_TVD3, 3, grid=grid)
        assert np.array_equal(Ug, ref) and abs(rg - r0) <= 1e-14 * abs(r0), grid
    # one launch per order bucket, the buckets side by side on the side streams:
    # compiled apart from the persistent launch, so equal to FP64 rounding
    Ug, rg = steps_under(g, U, dt, po.RK_TVD3, 3, shape=1)
    assert rel(Ug, ref) < 1e-12 and abs(rg - r0) <= 1e-12 * abs(r0)


def test_autotune_restores_state_and_uses_cache(tmp_path):
    announce("autotune + cache")
    p = 3
    o, g = tri_pair(p)
    U = smooth_euler(o, p, o.ndof // (4 * np_tri(p)), 9)
    g.set_state(U)
    dt = o.dt(U, 0.1)
    cache = str(tmp_path / "tune.txt")
    rep = g.autotune(po.RK4, dt, steps=3, cache=cache)
    assert not rep["from_cache"]
    assert [c[:2] for c in rep["candidates"]] == [(0, 0), (1, 0)]
    assert all(c[2] > 0 for c in rep["candidates"])
    assert rep["best"] in [c[:2] for c in rep["candidates"]]
    assert g.schedule() == rep["best"]
    assert np.array_equal(g.get_state(), U)
    lines = open(cache).read().splitlines()
    assert lines == ["%s\t%d %d" % (rep["key"], rep["best"][0], rep["best"][1])]
    assert rep["key"].split()[0].startswith("gfx")
    # a second context of the same configuration: from the cache, nothing timed
    o2, g2 = tri_pair(p)
    g2.set_state(U)
    rep2 = g2.autotune(po.RK4, dt, steps=3, cache=cache)
    assert rep2["from_cache"] and rep2["best"] == rep["best"] and rep2["key"] == rep["key"]
    assert g2.schedule() == rep["best"]
    # forced: timed again and appended
    rep3 = g2.autotune(po.RK4, dt, steps=3, cache=cache, force=True)
    assert not rep3["from_cache"] and len(open(cache).read().splitlines()) == 2
    # the tuned context steps exactly as an untuned one
    Ua, _ = steps_under(g2, U, dt, po.RK4, 2, *g2.schedule())
    g.set_state(U)
    g.set_schedule(0, 0)
    g.advanceInTime(0.0, dt, scheme=po.RK4, nsteps=2)
    assert np.array_equal(Ua, g.get_state())


def test_autotune_mixed_and_limiter_flags_restored():
    announce("autotune mixed")
    g = mixed_ctx()
    g.project(2, (0.53,))
    g.set_limiter(True)
    flags = g.limit()
    U = g.get_state()
    dt = 0.5 * g.computeDt(0.1)
    rep = g.autotune(po.RK_TVD3, dt, steps=2)
    cand = [tuple(c[:2]) for c in rep["candidates"]]
    assert cand[0] == (0, 0) and cand[-1] == (1, 0) and len(cand) >= 3
    assert all(c[0] == 0 for c in cand[:-1])
    assert np.array_equal(g.get_state(), U)
    assert np.array_equal(g.limiter_flags(), flags)


def test_bad_schedules_refused():
    import dgpar
    announce("bad schedules")
    o, g = tri_pair(2)
    with pytest.raises(dgpar.DGError):
        g.set_schedule(shape=0, grid=4)   # no persistent launch on a uniform mesh
    with pytest.raises(dgpar.DGError):
        g.set_schedule(shape=2)
    gm = mixed_ctx()
    with pytest.raises(dgpar.DGError):
        gm.set_schedule(shape=2)          # mixed orders: persistent (0) or bucket by bucket (1)
    with pytest.raises(dgpar.DGError):
        gm.set_schedule(shape=1, grid=4)  # the bucket launches have no persistent grid
    with pytest.raises(dgpar.DGError):
        gm.set_schedule(grid=10 ** 6)     # more workgroups than chunks
    assert g.schedule() == (0, 0) and gm.schedule() == (0, 0)


@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_tri_mfma_tiles_fp32_form(p):
    """shape 2 (FP32 form of the Euler law): the test contraction and the face
    lift on the matrix cores; the residual follows the FP64 oracle to the FP32
    form's budget and shape 0 to FP32 rounding, and RK4 steps follow shape 0"""
    import dgpar
    announce("mfma tiles P%d" % p)
    o, g = tri_pair(p, n=9, precision=dgpar.PREC_FP32)  # 162 elements: a partial last tile
    U = smooth_euler(o, p, o.ndof // (4 * np_tri(p)), 13)
    g.set_state(U)
    R0 = g.residual()
    g.set_schedule(shape=2)
    assert g.schedule() == (2, 0)
    R2 = g.residual()
    Ro = o.residual(U)[0]
    assert np.abs(R2 - Ro).max() <= 5e-5 * np.abs(Ro).max()
    assert np.abs(R2 - R0).max() <= 2e-6 * np.abs(Ro).max()
    dt = o.dt(U, 0.1)
    U0, r0 = steps_under(g, U, dt, po.RK4, 5, shape=0)
    U2, r2 = steps_under(g, U, dt, po.RK4, 5, shape=2)
    assert np.abs(U2 - U0).max() <= 1e-6 * np.abs(U0).max()
    assert abs(r2 - r0) <= 1e-4 * abs(r0)


def test_autotune_fp32_offers_the_mfma_shape():
    import dgpar
    announce("autotune fp32 candidates")
    o, g = tri_pair(3, precision=dgpar.PREC_FP32)
    U = smooth_euler(o, 3, o.ndof // (4 * np_tri(3)), 9)
    g.set_state(U)
    rep = g.autotune(po.RK4, o.dt(U, 0.1), steps=3)
    assert [c[:2] for c in rep["candidates"]] == [(0, 0), (1, 0), (2, 0)]
    assert np.array_equal(g.get_state(), U)
