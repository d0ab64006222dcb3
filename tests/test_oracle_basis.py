"""Pins the oracle's own Gram-Schmidt basis (and the generated tables the product
uses) by orthonormality, the triangular Gram-Schmidt structure, and the
reference's printed tables read as text (OrthoSFTri.h, TetOrtho.h)."""
import os
import re

import numpy as np
import pytest

from helpers import REFERENCE, collapsed_rule
from oracle import pyoracle


def tri_monomials(n):
    out, d = [], 0
    while len(out) < n:
        out += [(d - i, i) for i in range(d + 1)]
        d += 1
    return out[:n]


@pytest.mark.parametrize("p", [0, 1, 2, 3, 4, 5])
def test_tri_orthonormal(p):
    pts, w = collapsed_rule(p + 2)
    vals = np.array([pyoracle.basis(2, p, (u, v, 0))[0] for u, v in pts])
    M = (vals * w[:, None]).T @ vals
    assert np.abs(M - np.eye(M.shape[0])).max() < (1e-13 if p <= 4 else 1e-12)


@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_tri_gram_schmidt_structure(p):
    # phi_i lies in span(m_0..m_i) with a positive coefficient on m_i
    n = (p + 1) * (p + 2) // 2
    rng = np.random.default_rng(1)
    pts = rng.random((200, 2))
    pts = pts[pts.sum(1) < 1]
    mon = tri_monomials(n)
    A = np.array([[u ** a * v ** b for a, b in mon] for u, v in pts])
    vals = np.array([pyoracle.basis(2, p, (u, v, 0))[0] for u, v in pts])
    for i in range(n):
        c, *_ = np.linalg.lstsq(A[:, : i + 1], vals[:, i], rcond=None)
        assert np.abs(A[:, : i + 1] @ c - vals[:, i]).max() < 1e-9
        assert c[i] > 0


def test_tri_gradients_match_finite_differences():
    x = np.array([0.21, 0.33, 0.0])
    h = 1e-6
    phi, dphi = pyoracle.basis(2, 4, x)
    for d in range(2):
        xp, xm = x.copy(), x.copy()
        xp[d] += h
        xm[d] -= h
        fd = (pyoracle.basis(2, 4, xp)[0] - pyoracle.basis(2, 4, xm)[0]) / (2 * h)
        assert np.abs(fd - dphi[:, d]).max() < 1e-6


def test_tet_orthonormal():
    # tensor collapsed rule on the tetrahedron
    x, w = np.polynomial.legendre.leggauss(6)
    t, wt = 0.5 * (x + 1), 0.5 * w
    pts, wts = [], []
    for a, wa in zip(t, wt):
        for b, wb in zip(t, wt):
            for c, wc in zip(t, wt):
                u = a
                v = (1 - a) * b
                z = (1 - a) * (1 - b) * c
                pts.append((u, v, z))
                wts.append(wa * wb * wc * (1 - a) ** 2 * (1 - b))
    wts = np.array(wts)
    vals = np.array([pyoracle.basis(3, 3, p)[0] for p in pts])
    M = (vals * wts[:, None]).T @ vals
    assert np.abs(M - np.eye(20)).max() < 1e-12


def _read_table(path, name):
    txt = open(path).read()
    body = txt[txt.index(name):]
    body = body[body.index("{") + 1:]
    nums = re.findall(r"[-+]?\d+\.\d+E[-+]\d+", body)
    return nums


@pytest.mark.skipif(not os.path.exists(os.path.join(REFERENCE, "OrthoSFTri.h")), reason="reference not mounted")
def test_tri_basis_matches_reference_table():
    nums = np.array([float(s) for s in _read_table(os.path.join(REFERENCE, "OrthoSFTri.h"), "myGramSchmidtTriangle[")])
    C = nums[: 45 * 45].reshape(45, 45)
    rng = np.random.default_rng(3)
    pts = rng.random((50, 2))
    pts = pts[pts.sum(1) < 1]
    mon = tri_monomials(15)
    for u, v in pts:
        m = np.array([u ** a * v ** b for a, b in mon])
        ref = C[:15, :15] @ m
        ours = pyoracle.basis(2, 4, (u, v, 0))[0]
        # P0..P2 rows agree to ~1e-13; the printed P3 rows differ from exact
        # Gram-Schmidt by a few 1e-12 and the P4 rows carry ~1e-9 noise
        assert np.abs(ref[:6] - ours[:6]).max() < 5e-13
        assert np.abs(ref[6:10] - ours[6:10]).max() < 2e-11
        assert np.abs(ref[10:15] - ours[10:15]).max() < 1e-7


@pytest.mark.skipif(not os.path.exists(os.path.join(REFERENCE, "TetOrtho.h")), reason="reference not mounted")
def test_tet_basis_matches_reference_table():
    nums = _read_table(os.path.join(REFERENCE, "TetOrtho.h"), "myTetOrtho[")
    nums = np.array([float(s) for s in nums])
    n = int(round(np.sqrt(len(nums))))
    C = nums[: n * n].reshape(n, n)
    mon = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (2, 0, 0), (0, 2, 0), (0, 0, 2), (1, 1, 0), (1, 0, 1),
           (0, 1, 1), (3, 0, 0), (0, 3, 0), (0, 0, 3), (2, 1, 0), (2, 0, 1), (1, 2, 0), (0, 2, 1), (1, 0, 2),
           (0, 1, 2), (1, 1, 1)]
    rng = np.random.default_rng(4)
    for _ in range(20):
        p = rng.random(3) / 3
        m = np.array([p[0] ** a * p[1] ** b * p[2] ** c for a, b, c in mon])
        ref = C[:20, :20] @ m
        ours = pyoracle.basis(3, 3, p)[0]
        assert np.abs(ref - ours).max() < 1e-7 * max(1.0, np.abs(ref).max())
