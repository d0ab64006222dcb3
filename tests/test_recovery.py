"""Pins the oracle's non-physical recovery (DGCell::setToZero, DGCell.cc:251-279:
a cell whose Euler state has rho < 0 or p < 0 at a point the residual
evaluates keeps its means and drops every mode >= 1):

- against the reference's own assembly (oracle/_ref): a cell made non-physical
  at its volume points only (a negative bubble lambda_1 lambda_2 lambda_3 in
  the density, zero on every edge) is reduced by the reference inside its
  volume pass (DGCell.cc:198-223); its residual equals the oracle's residual of
  the recovered state;
- physical states are untouched, means never change, a second pass changes
  nothing, flagged elements keep only their means;
- an element non-physical only at points of its faces is found: the density
  trace, evaluated independently from the basis at the Gauss-Legendre points of
  that face, is negative there while every volume point stays positive."""
import os

import numpy as np
import pytest

import ref_harness as rh
from helpers import REFERENCE, collapsed_rule, np_tri
from oracle import pyoracle as po

FF = (1.0, 0.2, -0.1, 2.6)


def bubble_coefficients(p):
    """modal coefficients of lambda_1 lambda_2 lambda_3 = xi eta (1 - xi - eta) on
    the reference triangle (exact: it lies in P3; symmetric in the vertices, so
    the same in every vertex order)"""
    pts, w = collapsed_rule(8)
    c = np.zeros(np_tri(p))
    for (x, y), wq in zip(pts, w):
        phi, _ = po.basis(2, p, (x, y, 0))
        c += wq * x * y * (1 - x - y) * phi
    return c


def free_stream_state(E, p):
    U = np.zeros((E, 4, np_tri(p)))
    U[:, :, 0] = np.array(FF) / np.sqrt(2.0)
    return U


@pytest.mark.skipif(not (rh.available() and os.path.exists(REFERENCE)), reason="reference build oracle/_ref not available")
@pytest.mark.parametrize("p", [3, 4])
def test_recovery_matches_the_reference_volume_pass(p):
    m = rh.square_mesh(5, jitter=0.12, seed=21, periodic=True)
    o = po.Oracle(2, m["top"], m["xyz"], p, po.LAW_EULER2D, po.FLUX_LF, gamma=1.4, farfield=FF, bc=m["bc"])
    U = free_stream_state(o.E, p)
    U += 0.002 * np.random.default_rng(p).standard_normal(U.shape) * (np.arange(np_tri(p)) > 0)
    b = bubble_coefficients(p)
    # the bubble is 1/27 at the centroid and averages 1/60: with a = -40 the density
    # is negative inside (-0.48 at the centroid), its mean stays positive (0.33)
    # and on the edges it is unchanged
    U[7, 0, :] += -40.0 * FF[0] * b
    U[12, 0, :] += -10.0 * FF[0] * b   # stays positive: untouched
    U = U.ravel()
    Ur, flags = o.recover(U)
    assert flags[7] == 1 and flags.sum() == 1
    Rr, _, _ = rh.run(m, "euler", "lf", p, "residual", U, farfield=FF)
    Ro, _ = o.residual(Ur)
    tol = {3: 2e-11, 4: 5e-9}[p]  # the reference's printed tables (as the residual tests)
    assert np.abs(Ro - Rr).max() / np.abs(Rr).max() < tol
    # the unrecovered residual differs from the reference's at that cell
    Rn, _ = o.residual(U)
    assert np.abs(Rn - Rr).max() / np.abs(Rr).max() > 1e-3


def tube_oracle(p=3, n=8):
    from helpers import mesh_tri
    vids, xyz, bc = mesh_tri(n, n, lx=1.0, ly=1.0, periodic_x=False, periodic_y=False, jitter=0.1, seed=4,
                             side_tags=(40000, 1, 40000, 1))
    orders = p if np.isscalar(p) else p
    return po.Oracle(2, vids, xyz, orders, po.LAW_EULER2D, po.FLUX_ROE_EF, gamma=1.4, farfield=FF, bc=bc), vids, xyz


def test_recovery_leaves_physical_states_and_is_idempotent():
    o, _, _ = tube_oracle(3)
    U = free_stream_state(o.E, 3)
    U += 0.01 * np.random.default_rng(1).standard_normal(U.shape) * (np.arange(10) > 0)
    U = U.ravel()
    V, f = o.recover(U)
    assert not f.any() and np.array_equal(V, U)
    # a few elements pushed negative in density through their high modes
    W = U.reshape(o.E, 4, 10).copy()
    rng = np.random.default_rng(2)
    bad = rng.choice(o.E, 6, replace=False)
    W[bad, 0, 1:] = rng.standard_normal((6, 9)) * 2.0
    W = W.ravel()
    X, f = o.recover(W)
    Xr = X.reshape(o.E, 4, 10)
    assert set(np.nonzero(f)[0]) <= set(bad) and (f[bad] & 1).sum() >= 5
    red = np.nonzero(f & 1)[0]
    assert (Xr[red, :, 1:] == 0).all()
    assert np.array_equal(Xr[:, :, 0], W.reshape(o.E, 4, 10)[:, :, 0])   # means kept
    keep = np.setdiff1d(np.arange(o.E), red)
    assert np.array_equal(Xr[keep], W.reshape(o.E, 4, 10)[keep])
    Y, f2 = o.recover(X)
    assert np.array_equal(Y, X) and not f2.any()


def test_recovery_flags_a_mean_and_nonfinite_coefficients():
    o, _, _ = tube_oracle(2)
    U = free_stream_state(o.E, 2)
    U[3, 0, 0] = -0.1     # negative mean density
    U[5, 2, 4] = np.nan
    X, f = o.recover(U.ravel())
    assert f[3] == 3      # reduced, mean non-physical
    assert f[5] & 4
    assert np.count_nonzero(f) == 2


def test_recovery_finds_a_face_only_violation():
    """density linear in the element, lowest at vertex 0: below zero at the
    Gauss-Legendre points of the edges next to vertex 0, positive at every
    volume point of the degree-2p rule"""
    from helpers import mesh_tri
    p = 1
    vids, xyz, bc = mesh_tri(4, 4, periodic_x=False, periodic_y=False, jitter=0.0, side_tags=(40000,) * 4)
    o = po.Oracle(2, vids, xyz, p, po.LAW_EULER2D, po.FLUX_LF, gamma=1.4, farfield=FF, bc=bc)
    U = free_stream_state(o.E, p)
    e = 9
    # rho(xi, eta) = a + b (xi + eta) in the element's reference frame
    pts = [(0.0, 0.0), (1.0, 0.0), (0.0, 1.0)]
    B = np.array([po.basis(2, p, (x, y, 0))[0] for x, y in pts])
    # rho = -0.4 + 1.4 (xi + eta): -0.10 at the edges' Gauss-Legendre points next to
    # vertex 0 (t = 0.211), +0.07 at the nearest point of a degree-2 rule
    target = np.array([-0.4, 1.0, 1.0])   # values at the vertices: rho < 0 only near vertex 0
    U[e, 0, :] = np.linalg.solve(B, target)
    # pressure kept positive where rho > 0: energy follows rho
    U[e, 3, :] = U[e, 0, :] * FF[3] / FF[0]
    U[e, 1, :] = 0
    U[e, 2, :] = 0
    U = U.ravel()
    # independent evaluation at the edges' Gauss-Legendre points and at the points
    # of both degree-2 rules (Strang-Fix and edge midpoints)
    gl = 0.5 * (np.polynomial.legendre.leggauss(2)[0] + 1)
    edge_pts = [(t, 0.0) for t in gl] + [(1 - t, t) for t in gl] + [(0.0, t) for t in gl]
    rho = lambda x, y: np.dot(po.basis(2, p, (x, y, 0))[0], U.reshape(o.E, 4, 3)[e, 0])  # noqa: E731
    assert min(rho(x, y) for x, y in edge_pts) < 0
    vol = [(1 / 6, 1 / 6), (2 / 3, 1 / 6), (1 / 6, 2 / 3), (0.5, 0.0), (0.5, 0.5), (0.0, 0.5)]
    assert min(rho(x, y) for x, y in vol) > 0
    _, f = o.recover(U)
    assert f[e] == 1 and np.count_nonzero(f) == 1
