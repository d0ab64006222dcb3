"""Mixed orders P1-P4, the triangle limiter and the order sensor.

Oracle properties (what pins the new limiter and sensor, which have no
triangle counterpart in the reference):
- a constant state is left alone, a linear state is left alone on a lattice of
  parallelograms (each neighbour's centroid is the mirror image of the
  element's across the shared edge midpoint), cell means never change, and a
  limited field keeps its edge-midpoint values between its mean and the
  neighbour means (nu = 1), so a second pass changes nothing beyond rounding;
- the sensor keeps constant elements, lowers an element whose energy sits in
  its top-degree modes, raises one with none there, and respects pmin/pmax.

Kernel functions run through the CPU emulation (tests/cpp/emu_stage.cc)
against the oracle: the mixed-order residual and limited SSP-RK3 steps within
1e-12, limiter flags and sensor orders bit for bit on the same state."""
import ctypes

import numpy as np
import pytest

from helpers import mesh_tri, np_tri
from oracle import pyoracle as po
from test_kernel_emulation import LIBEMU

SOD_FF = (1.0, 0.0, 0.0, 2.5, 0.125, 0.0, 0.0, 0.25)  # left state, right state (TAG 1, TAG 2)
PD = ctypes.POINTER(ctypes.c_double)
PU8 = ctypes.POINTER(ctypes.c_uint8)
PI32 = ctypes.POINTER(ctypes.c_int32)


def sod_case(orders_seed=None, p=2, nx=16, ny=4, jitter=0.1, flux=po.FLUX_ROE_EF, x0=0.53):
    """shock tube on [0,1]x[0,0.25]: walls at the bottom and top, the left and
    right states as far field at the two ends"""
    vids, xyz, bc = mesh_tri(nx, ny, lx=1.0, ly=0.25, periodic_x=False, periodic_y=False, jitter=jitter, seed=4,
                             side_tags=(40000, 2, 40000, 1))
    E = vids.shape[0]
    orders = (np.random.default_rng(orders_seed).integers(1, 5, E).astype(np.int32) if orders_seed is not None
              else np.full(E, p, np.int32))
    o = po.Oracle(2, vids, xyz, orders, po.LAW_EULER2D, flux, gamma=1.4, farfield=SOD_FF, bc=bc)
    return vids, xyz, bc, orders, o, o.project(2, (x0,))


def blocks(U, orders, nf=4):
    sizes = nf * (orders + 1) * (orders + 2) // 2
    offs = np.concatenate([[0], np.cumsum(sizes)])
    return [U[offs[e]:offs[e + 1]].reshape(nf, -1) for e in range(len(orders))]


def test_limiter_leaves_constants_alone():
    vids, xyz, bc, orders, o, U = sod_case(3)
    U = np.concatenate([np.pad(np.array([SOD_FF[:4]]).T / np.sqrt(0.5) * 0.5, ((0, 0), (0, np_tri(p) - 1))).ravel()
                        for p in orders])
    Ul, fl = o.limit(U)
    assert not fl.any()
    assert np.array_equal(Ul, U)


@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_limiter_leaves_linear_fields_alone_on_parallelograms(p):
    vids, xyz, bc = mesh_tri(8, 8, lx=1.0, ly=1.0, periodic_x=False, periodic_y=False, jitter=0.0, seed=1,
                             side_tags=(7, 7, 7, 7))
    o = po.Oracle(2, vids, xyz, p, po.LAW_SCALAR, po.FLUX_UPWIND, vel=(1.0, 0.5), farfield=(0.0,), bc=bc)
    # u = 1 + 0.3 x - 0.7 y in each element's modes (the linear modes fitted exactly)
    U = np.zeros(o.ndof)
    npp = np_tri(p)
    for e in range(o.E):
        X = xyz[e]
        # u(xi) = u(x0) + grad u . J xi, J columns = edges
        g = np.array([0.3, -0.7])
        J = np.array([X[1] - X[0], X[2] - X[0]]).T
        gr = J.T @ g  # d u / d xi
        u0 = 1 + g @ X[0]
        # mean over the reference triangle and the linear modes from a fit at 3 points
        pts = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1 / 3, 1 / 3, 0]])
        vals = u0 + pts[:, :2] @ gr
        Phi = np.array([po.basis(2, 1, x)[0] for x in pts])
        c = np.linalg.lstsq(Phi, vals, rcond=None)[0]
        U[e * npp:e * npp + 3] = c
    Ul, fl = o.limit(U)
    faces = o.faces()
    bnd = np.zeros(o.E, bool)  # elements with a boundary face (ghost means are not linear data)
    bnd[faces[faces[:, 2] < 0, 0]] = True
    assert not fl[~bnd].any()
    for e in np.where(~bnd)[0]:
        assert np.array_equal(Ul[e * npp:(e + 1) * npp], U[e * npp:(e + 1) * npp])


@pytest.mark.parametrize("seed", [None, 5])
def test_limiter_keeps_means_and_bounds_midpoint_values(seed):
    vids, xyz, bc, orders, o, U = sod_case(seed, p=3)
    Ul, fl = o.limit(U)
    assert fl.any()
    B0, B1 = blocks(U, orders), blocks(Ul, orders)
    phi0 = po.basis(2, 0, [0, 0, 0])[0][0]
    for e in range(o.E):
        assert np.array_equal(B0[e][:, 0], B1[e][:, 0])  # means untouched, bit for bit
    # limited fields: midpoint values between the mean and the neighbour means
    faces = o.faces()
    nbr = {e: {} for e in range(o.E)}
    for eL, lL, eR, lR in faces:
        if eR >= 0:
            nbr[eL][lL] = eR
            nbr[eR][lR] = eL
    mids = np.array([[0.5, 0, 0], [0.5, 0.5, 0], [0, 0.5, 0]])
    checked = 0
    for e in np.where(fl > 0)[0]:
        p = orders[e]
        for f in range(4):
            if not fl[e] & (1 << f):
                continue
            m = B1[e][f, 0] * phi0
            for l, c in enumerate(mids):
                if l not in nbr[e]:
                    continue
                mn = B1[nbr[e][l]][f, 0] * phi0
                u = B1[e][f] @ po.basis(2, p, c)[0]
                assert min(m, mn) - 1e-13 <= u <= max(m, mn) + 1e-13
                checked += 1
    assert checked > 0
    Ul2, _ = o.limit(Ul)
    assert np.abs(Ul2 - Ul).max() <= 1e-14 * np.abs(Ul).max()


def test_sensor_rules():
    vids, xyz, bc, orders, o, U = sod_case(None, p=3)
    npp = np_tri(3)
    V = np.zeros((o.E, 4, npp))
    V[:, :, 0] = 1.0
    V[0, 0, 7] = 0.5      # element 0: energy in a degree-3 mode -> lower
    V[1, 0, 1] = 0.5      # element 1: energy only below degree 3 -> raise
    V[2, 0, 1] = 0.5      # element 2: both, top share 0.09/1.34 between the thresholds -> keep
    V[2, 0, 6] = 0.3
    s_hi, s_lo = [0.1] * 4, [0.01] * 4
    got = o.sensor(V.ravel(), s_hi, s_lo)
    assert got[0] == 2 and got[1] == 4 and got[2] == 3
    assert (got[3:] == 3).all()  # constant elements keep their order
    got = o.sensor(V.ravel(), s_hi, s_lo, pmin=3, pmax=3)
    assert (got == 3).all()


# ---------------------------------------------------------------------------
# emulated kernels against the oracle

pytestmark_emu = pytest.mark.skipif(LIBEMU is None, reason="emulator not built")
if LIBEMU is not None:
    import dgpar

    p_md, p_ld = ctypes.POINTER(dgpar.MeshDesc), ctypes.POINTER(dgpar.LawDesc)
    LIBEMU.emu_mixed_residual.argtypes = [p_md, p_ld, PD, PD]
    LIBEMU.emu_mixed_steps.argtypes = [p_md, p_ld, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_double, ctypes.c_double, PD, PD, PU8]
    LIBEMU.emu_limit.argtypes = [p_md, p_ld, ctypes.c_int, ctypes.c_double, ctypes.c_double, PD, PU8]
    LIBEMU.emu_sensor.argtypes = [p_md, p_ld, ctypes.c_int, PD, PD, PD, ctypes.c_int, ctypes.c_int, PI32]


def descs(vids, xyz, bc, orders, flux, ff=SOD_FF, law="euler"):
    mesh = dgpar.Mesh(2, vids, xyz, bc, order=orders)
    if law == "euler":
        lw = dgpar.Law(dgpar.LAW_EULER2D, flux, gamma=1.4, farfield=ff)
    else:
        lw = dgpar.Law(dgpar.LAW_SCALAR, dgpar.FLUX_UPWIND, vel=(1.0, 0.5), farfield=ff)
    return mesh, lw


@pytestmark_emu
@pytest.mark.parametrize("flux", [po.FLUX_ROE, po.FLUX_LF, po.FLUX_ROE_EF])
@pytest.mark.parametrize("seed", [1, 2])
def test_emulated_mixed_residual(flux, seed):
    vids, xyz, bc, orders, o, _ = sod_case(seed, flux=flux)
    # a smooth physical state: the limited tube with 1% noise on every mode
    rng = np.random.default_rng(seed)
    blocks_ = []
    for p in orders:
        b = 0.01 * rng.standard_normal((4, np_tri(p)))
        b[:, 0] += np.array([0.6, 0.1, -0.05, 1.4]) / np.sqrt(2.0)
        blocks_.append(b.ravel())
    U = np.concatenate(blocks_)
    mesh, lw = descs(vids, xyz, bc, orders, flux)
    md, ld = mesh.desc(), lw.desc()
    R = np.zeros_like(U)
    assert LIBEMU.emu_mixed_residual(ctypes.byref(md), ctypes.byref(ld), U.ctypes.data_as(PD), R.ctypes.data_as(PD)) == 0
    Ro, _ = o.residual(U)
    assert np.abs(R - Ro).max() <= 1e-12 * np.abs(Ro).max()


@pytestmark_emu
@pytest.mark.parametrize("mode", [0, 1])
def test_emulated_mixed_residual_many_tiles(mode):
    """about 180 elements per order bucket: whole tiles of 64 and a partial last
    tile in every bucket; mode 0 the tiles of the persistent launch (LDS rows of
    the top order), mode 1 those of the bucket-by-bucket launches (each bucket's
    own LDS rows)"""
    LIBEMU.emu_mixed_mode(mode)
    vids, xyz, bc, orders, o, _ = sod_case(3, nx=36, ny=10)
    assert all(np.bincount(orders, minlength=5)[1:] > 128)
    rng = np.random.default_rng(3)
    U = np.concatenate([(0.01 * rng.standard_normal((4, np_tri(p))) +
                         np.outer([0.6, 0.1, -0.05, 1.4], np.eye(np_tri(p))[0]) / np.sqrt(2.0)).ravel()
                        for p in orders])
    mesh, lw = descs(vids, xyz, bc, orders, po.FLUX_ROE_EF)
    md, ld = mesh.desc(), lw.desc()
    R = np.zeros_like(U)
    try:
        assert LIBEMU.emu_mixed_residual(ctypes.byref(md), ctypes.byref(ld), U.ctypes.data_as(PD),
                                         R.ctypes.data_as(PD)) == 0
    finally:
        LIBEMU.emu_mixed_mode(0)
    Ro, _ = o.residual(U)
    assert np.abs(R - Ro).max() <= 1e-12 * np.abs(Ro).max()


@pytestmark_emu
def test_emulated_mixed_scalar_residual():
    vids, xyz, bc = mesh_tri(6, 6, periodic_x=True, periodic_y=False, jitter=0.1, seed=2, side_tags=(7, 7, 7, 7))
    orders = np.random.default_rng(8).integers(1, 5, vids.shape[0]).astype(np.int32)
    o = po.Oracle(2, vids, xyz, orders, po.LAW_SCALAR, po.FLUX_UPWIND, vel=(1.0, 0.5), farfield=(0.2,), bc=bc)
    U = np.random.default_rng(3).standard_normal(o.ndof)
    mesh, lw = descs(vids, xyz, bc, orders, 0, ff=(0.2,), law="scalar")
    md, ld = mesh.desc(), lw.desc()
    R = np.zeros_like(U)
    assert LIBEMU.emu_mixed_residual(ctypes.byref(md), ctypes.byref(ld), U.ctypes.data_as(PD), R.ctypes.data_as(PD)) == 0
    Ro, _ = o.residual(U)
    assert np.abs(R - Ro).max() <= 1e-12 * np.abs(Ro).max()


@pytestmark_emu
@pytest.mark.parametrize("seed", [None, 6])
def test_emulated_limiter_and_sensor_bit_exact(seed):
    vids, xyz, bc, orders, o, U = sod_case(seed, p=3)
    mesh, lw = descs(vids, xyz, bc, orders, po.FLUX_ROE_EF)
    md, ld = mesh.desc(), lw.desc()
    for M, nu in ((0.0, 1.0), (5.0, 1.5)):
        Ue = U.copy()
        fe = np.zeros(o.E, np.uint8)
        assert LIBEMU.emu_limit(ctypes.byref(md), ctypes.byref(ld), 3, M, nu, Ue.ctypes.data_as(PD),
                                fe.ctypes.data_as(PU8)) == 0
        Uo, fo = o.limit(U, M, nu)
        assert np.array_equal(fe, fo)
        assert np.abs(Ue - Uo).max() <= 1e-14 * np.abs(Uo).max()
    s_hi = np.array([0.05, 0.02, 0.01, 0.005])
    s_lo = s_hi * 1e-3
    oe = np.zeros(o.E, np.int32)
    assert LIBEMU.emu_sensor(ctypes.byref(md), ctypes.byref(ld), 3, U.ctypes.data_as(PD), s_hi.ctypes.data_as(PD),
                             s_lo.ctypes.data_as(PD), 1, 4, oe.ctypes.data_as(PI32)) == 0
    assert np.array_equal(oe, o.sensor(U, s_hi, s_lo, 1, 4))


@pytestmark_emu
@pytest.mark.parametrize("seed,scheme", [(7, po.RK_TVD3), (None, po.RK_TVD3), (9, po.RK4)])
def test_emulated_limited_mixed_steps(seed, scheme):
    vids, xyz, bc, orders, o, U = sod_case(seed, p=2)
    U, _ = o.limit(U)
    dt = o.dt(U, 0.1)
    o.set_limiter(True)
    Uo, t = U.copy(), 0.0
    for _ in range(4):
        Uo, ro = o.step(Uo, scheme, t, dt)
        t += dt
    _, fo = o.limit(Uo)  # flags of a pass over the final state (unchanged by it)
    mesh, lw = descs(vids, xyz, bc, orders, po.FLUX_ROE_EF)
    md, ld = mesh.desc(), lw.desc()
    Ue = U.copy()
    r = ctypes.c_double()
    fe = np.zeros(o.E, np.uint8)
    assert LIBEMU.emu_mixed_steps(ctypes.byref(md), ctypes.byref(ld), scheme, dt, 4, 1, 0.0, 1.0,
                                  Ue.ctypes.data_as(PD), ctypes.byref(r), fe.ctypes.data_as(PU8)) == 0
    assert np.abs(Ue - Uo).max() <= 1e-12 * np.abs(Uo).max()
    assert abs(r.value - ro) <= 1e-10 * abs(ro)


def test_positivity_scaling_keeps_means_and_point_values_positive():
    """Zhang-Shu scaling: after the limiter, density and pressure are positive at
    the volume points, the vertices and the edge points of every face rule, and
    the means are untouched"""
    vids, xyz, bc, orders, o, U = sod_case(21, p=4)
    rng = np.random.default_rng(2)
    B = blocks(U, orders)
    for b in B:  # overshooting high modes everywhere
        b[:, 1:] += 0.3 * rng.standard_normal(b[:, 1:].shape) * np.abs(b[:, :1])
    U = np.concatenate([b.ravel() for b in B])
    Ul, fl = o.limit(U, 0.0, 1.0)
    assert (fl & 16).any()
    B1 = blocks(Ul, orders)
    for b0, b1 in zip(blocks(U, orders), B1):
        assert np.array_equal(b0[:, 0], b1[:, 0])
    gl = {n: np.polynomial.legendre.leggauss(n)[0] for n in range(2, 6)}
    V = np.array([[0, 0], [1, 0], [0, 1]])
    for e, b in enumerate(B1):
        p = orders[e]
        pts = [V[0], V[1], V[2]]
        for n in range(p + 1, 6):
            for l in range(3):
                for s in gl[n]:
                    pts.append(0.5 * (1 + s) * V[l] + 0.5 * (1 - s) * V[(l + 1) % 3])
        for x in pts:
            w = b @ po.basis(2, p, [x[0], x[1], 0.0])[0]
            assert w[0] > 0
            assert 0.4 * (w[3] - 0.5 * (w[1] ** 2 + w[2] ** 2) / w[0]) > -1e-12
