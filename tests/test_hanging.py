"""Meshes with hanging nodes (cells split by red refinement beside unsplit
neighbours, the reference's mDGMesh::split, mDGMesh.cc:54-337): the oracle's
faces of half an edge (child left, the coarse cell right, both evaluated at the
half's Gauss points, as the reference's boundary cells of leaf edges do,
DGCell.cc:423-555) pinned by closed forms.

The reference's own split path could not serve as a check: a DGAnalysis built
on a mesh refined by mDGMesh::split stops with a null dereference in
DGBoundaryCell::init (GaussIntegrator::iPoint, Integrator.cc:58).  What pins
the oracle here:
  * a free stream gives a zero residual (normals, half-face weights, both sides);
  * every face flux enters both sides with opposite signs: the mean residuals
    sum to zero on a periodic mesh for any state;
  * for a continuous linear field and linear advection the residual of an
    interior cell is -(v . grad u) int phi_k = -(v . grad u) |J| / sqrt 2 for
    k = 0 and 0 above: the coarse side's trace at the half's points must
    match the child's exactly.
"""
import numpy as np
import pytest

from helpers import mesh_tri, refine_tri
from oracle import pyoracle as po

VORTEX = (5.0, 5.0, 5.0, 1.0, 1.0, 10.0)


def split_mesh(n=6, periodic=True, lx=1.0, pick=(7, 8, 21, 30, 31, 44), levels=1, seed=7):
    vids, xyz, bc = mesh_tri(n, n, lx=lx, ly=lx, periodic_x=periodic, periodic_y=periodic, jitter=0.1, seed=seed,
                             side_tags=(7, 7, 7, 7))
    mask = np.zeros(len(vids), bool)
    mask[list(pick)] = True
    v, x, b, par = refine_tri(vids, xyz, bc, mask)
    if levels == 2:
        # split two children of a split cell once more where their faces are on
        # split neighbours only (1-irregular)
        m2 = np.zeros(len(v), bool)
        m2[np.where(par == pick[0])[0][3]] = True  # the inner child
        v, x, b, par2 = refine_tri(v, x, b, m2)
    return v, x, b


def test_hanging_faces_are_found():
    v, x, b = split_mesh()
    o = po.Oracle(2, v, x, 1, po.LAW_SCALAR, po.FLUX_UPWIND, vel=(1.0, 0.5), bc=b)
    F = o.faces()
    assert (F[:, 2] >= 0).all()  # periodic: every face has two sides
    # 6 split cells: 4 children each; faces = 3 E / 2 + one extra per hanging face
    nh = len(F) - 3 * len(v) // 2
    assert nh > 0
    # the halves: child on the left, the coarse cell on the right
    per_cell = np.bincount(F[:, 2], minlength=len(v)) + np.bincount(F[:, 0], minlength=len(v))
    assert per_cell.max() >= 4  # a coarse cell with at least one face in two halves


@pytest.mark.parametrize("p", [1, 2, 3])
def test_free_stream(p):
    v, x, b = split_mesh(lx=10.0)
    o = po.Oracle(2, v, x, p, po.LAW_EULER2D, po.FLUX_ROE, gamma=1.4, bc=b)
    U = o.project(1, (0.0,) + VORTEX[1:])  # strength 0: the free stream (1, 1, 1, 1)
    R, _ = o.residual(U)
    # zero up to the rounding of the summed terms
    assert np.abs(R).max() <= 1e-14 * o.residual_scale(U).max()


@pytest.mark.parametrize("law,p", [("scalar", 2), ("euler", 1), ("euler", 3)])
def test_conservation(law, p):
    v, x, b = split_mesh(lx=10.0)
    if law == "scalar":
        o = po.Oracle(2, v, x, p, po.LAW_SCALAR, po.FLUX_UPWIND, vel=(1.0, 0.5), bc=b)
        U = o.project(0)
    else:
        o = po.Oracle(2, v, x, p, po.LAW_EULER2D, po.FLUX_LF, gamma=1.4, bc=b)
        U = o.project(1, VORTEX)
    U = U + 0.01 * np.random.default_rng(3).standard_normal(U.shape)  # discontinuous
    R, _ = o.residual(U)
    Np = (p + 1) * (p + 2) // 2
    R0 = R.reshape(len(v), -1, Np)[:, :, 0]
    scale = np.abs(o.residual_scale(U)).reshape(len(v), -1, Np)[:, :, 0].sum(axis=0)
    assert (np.abs(R0.sum(axis=0)) <= 1e-13 * scale).all()


@pytest.mark.parametrize("p,levels", [(1, 1), (2, 1), (3, 1), (2, 2)])
def test_linear_field_exact(p, levels):
    v, x, b = split_mesh(n=8, periodic=False, pick=(20, 21, 36, 52, 53, 70), levels=levels)
    vel = (1.0, 0.5)
    o = po.Oracle(2, v, x, p, po.LAW_SCALAR, po.FLUX_UPWIND, vel=vel, farfield=(0.0,), bc=b)
    a, gx, gy = 0.3, 0.7, -0.4
    U = o.project_fn(lambda X, Y, Z, t: (a + gx * X + gy * Y,))
    R, _ = o.residual(U)
    Np = (p + 1) * (p + 2) // 2
    R = R.reshape(len(v), Np)
    detj = o.detj()
    F = o.faces()
    boundary = np.zeros(len(v), bool)
    boundary[F[F[:, 2] < 0, 0]] = True
    expect = np.zeros_like(R)
    expect[:, 0] = -(vel[0] * gx + vel[1] * gy) * detj / np.sqrt(2.0)
    inner = ~boundary
    assert inner.sum() > len(v) // 2
    assert np.abs(R[inner] - expect[inner]).max() <= 1e-12 * np.abs(expect).max()


# ---- the stage kernel and the hanging-face pass, CPU emulation vs the oracle ---
def _emu():
    import ctypes
    import os

    import dgpar
    from helpers import ROOT
    path = os.path.join(ROOT, "tests", "cpp", "libemu.so")
    if not os.path.exists(path):
        pytest.skip("tests/cpp/libemu.so not built")
    lib = ctypes.CDLL(path)
    p_md, p_ld = ctypes.POINTER(dgpar.MeshDesc), ctypes.POINTER(dgpar.LawDesc)
    pd = ctypes.POINTER(ctypes.c_double)
    lib.emu_tri_residual.argtypes = [p_md, p_ld, ctypes.c_int, ctypes.c_int, pd, pd]
    lib.emu_tri_steps.argtypes = [p_md, p_ld, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                  ctypes.c_int, pd, pd]
    return lib, dgpar, ctypes


def hang_case(law, p, periodic=True):
    import dgpar
    v, x, b = split_mesh(lx=10.0, periodic=periodic)
    ff = (1.0, 0.1, 0.05, 2.5)
    if law == "scalar":
        o = po.Oracle(2, v, x, p, po.LAW_SCALAR, po.FLUX_UPWIND, vel=(1.0, 0.5), farfield=(0.2,), bc=b)
        lw = dgpar.Law(dgpar.LAW_SCALAR, dgpar.FLUX_UPWIND, vel=(1.0, 0.5), farfield=(0.2,))
        U = o.project_fn(lambda X, Y, Z, t: (np.exp(-((X - 5) ** 2 + (Y - 5) ** 2) / 4.0),))
    elif law == "maxwell2d":
        o = po.Oracle(2, v, x, p, po.LAW_MAXWELL2D, po.FLUX_UPWIND, farfield=(0.1, 0.0, 0.2), bc=b)
        lw = dgpar.Maxwell2dLaw(farfield=(0.1, 0.0, 0.2))
        U = o.project_fn(lambda X, Y, Z, t: (np.sin(0.6 * Y), np.cos(0.5 * X), np.sin(0.4 * (X + Y))))
    elif law == "burgers":
        o = po.Oracle(2, v, x, p, po.LAW_BURGERS2D, po.FLUX_CENTRAL, farfield=(0.4,), bc=b)
        lw = dgpar.Burgers2DLaw((0.4,))
        U = o.project_fn(lambda X, Y, Z, t: (0.5 + 0.3 * np.sin(0.6 * X) * np.cos(0.4 * Y),))
    else:
        fl = {"lf": (po.FLUX_LF, dgpar.FLUX_LF), "roe": (po.FLUX_ROE, dgpar.FLUX_ROE)}[law]
        o = po.Oracle(2, v, x, p, po.LAW_EULER2D, fl[0], gamma=1.4, farfield=ff, bc=b)
        lw = dgpar.Law(dgpar.LAW_EULER2D, fl[1], gamma=1.4, farfield=ff)
        U = o.project(1, VORTEX)
    U = U + 0.002 * np.random.default_rng(11).standard_normal(U.shape)
    return dgpar.Mesh(2, v, x, b), lw, o, U


@pytest.mark.parametrize("law,p,periodic", [("scalar", 1, True), ("scalar", 3, False), ("lf", 1, True),
                                             ("lf", 2, False), ("roe", 3, True), ("roe", 4, True),
                                             ("maxwell2d", 2, False), ("maxwell2d", 4, True), ("burgers", 3, True),
                                             ("burgers", 2, False)])
def test_emulated_residual_matches_oracle(law, p, periodic):
    lib, dgpar, ctypes = _emu()
    mesh, lw, o, U = hang_case(law, p, periodic)
    Ro, _ = o.residual(U)
    R = np.zeros_like(U)
    md, ld = mesh.desc(), lw.desc()
    pd = ctypes.POINTER(ctypes.c_double)
    assert lib.emu_tri_residual(ctypes.byref(md), ctypes.byref(ld), p, 0, U.ctypes.data_as(pd), R.ctypes.data_as(pd)) == 0
    assert np.abs(R - Ro).max() <= 1e-12 * np.abs(Ro).max()


@pytest.mark.parametrize("scheme", [po.RK4, po.RK_TVD3, po.RK_TVD2, po.RK_FE])
def test_emulated_steps_match_oracle(scheme):
    lib, dgpar, ctypes = _emu()
    mesh, lw, o, U = hang_case("lf", 2)
    dt = o.dt(U, 0.1)
    V = U
    for k in range(2):
        V, _ = o.step(V, scheme, k * dt, dt)
    W = np.array(U, copy=True)
    md, ld = mesh.desc(), lw.desc()
    pd = ctypes.POINTER(ctypes.c_double)
    r = ctypes.c_double()
    assert lib.emu_tri_steps(ctypes.byref(md), ctypes.byref(ld), 2, 0, scheme, dt, 2, W.ctypes.data_as(pd),
                             ctypes.byref(r)) == 0
    assert np.abs(W - V).max() <= 1e-12 * np.abs(V).max()
