"""Pins the generated quadrature tables (shared input of the product and the
oracle): exactness on every monomial up to the rule degree, weight sums, and
agreement with the reference's printed rules (GaussQuadratureTri.cc,
GaussQuadratureTet.cc, GaussLegendre1D.cc) read as text, as point sets."""
import math
import os
import re

import numpy as np
import pytest

from helpers import REFERENCE, ROOT

HDR = os.path.join(ROOT, "dg-parallel_amd", "csrc", "dg_rules_data.h")


def _array(name):
    txt = open(HDR).read()
    i = txt.index(name)
    body = txt[txt.index("{", i): txt.index("};", i)]
    return np.array([float(x) for x in re.findall(r"[-+]?\d+\.\d+e[-+]\d+|[-+]?\d+(?:\.\d+)?", body)])


def _ints(name):
    txt = open(HDR).read()
    i = txt.index(name)
    body = txt[txt.index("{", i): txt.index("}", i)]
    return [int(x) for x in re.findall(r"-?\d+", body)]


def tri_rules():
    off, npt = _ints("dg_tri_rule_off"), _ints("dg_tri_rule_npt")
    pts = _array("dg_tri_rule_pts[").reshape(-1, 3)
    return {d: pts[off[d]: off[d] + npt[d]] for d in range(len(off))}


def tet_rules():
    off, npt = _ints("dg_tet_rule_off"), _ints("dg_tet_rule_npt")
    pts = _array("dg_tet_rule_pts[").reshape(-1, 4)
    return {d: pts[off[d]: off[d] + npt[d]] for d in range(len(off))}


@pytest.mark.parametrize("deg", range(1, 11))
def test_tri_rule_exact(deg):
    r = tri_rules()[deg]
    assert abs(r[:, 2].sum() - 1.0) < 1e-15
    for n in range(deg + 1):
        for i in range(n + 1):
            a, b = n - i, i
            exact = 2.0 * math.factorial(a) * math.factorial(b) / math.factorial(a + b + 2)
            got = np.sum(r[:, 2] * r[:, 0] ** a * r[:, 1] ** b)
            assert abs(got - exact) < 2e-16 * 10, (a, b, got, exact)


@pytest.mark.parametrize("deg", [1, 2, 3, 4, 6])
def test_tet_rule_exact(deg):
    r = tet_rules()[deg]
    assert abs(r[:, 3].sum() - 1.0) < 1e-15
    for n in range(deg + 1):
        for i in range(n + 1):
            for j in range(n + 1 - i):
                a, b, c = i, j, n - i - j
                exact = 6.0 * math.factorial(a) * math.factorial(b) * math.factorial(c) / math.factorial(n + 3)
                got = np.sum(r[:, 3] * r[:, 0] ** a * r[:, 1] ** b * r[:, 2] ** c)
                assert abs(got - exact) < 2e-15


def test_gauss_legendre():
    off = _ints("dg_gl_off[12] =")
    pts = _array("dg_gl_pts[").reshape(-1, 2)
    for n in range(1, 13):
        r = pts[off[n - 1]: off[n - 1] + n]
        x, w = np.polynomial.legendre.leggauss(n)
        assert np.abs(np.sort(r[:, 0]) - x).max() < 1e-15
        assert np.abs(r[np.argsort(r[:, 0]), 1] - w).max() < 1e-14


def _ref_tri_rule(name, n):
    txt = open(os.path.join(REFERENCE, "GaussQuadratureTri.cc")).read()
    i = txt.index("IntPt2d %s[%d]" % (name, n))
    body = txt[txt.index("{", i): txt.index("};", i)]
    nums = [float(x) for x in re.findall(r"-?\d*\.\d+|-?\d+\.?", body)]
    return np.array(nums).reshape(n, 3)


@pytest.mark.skipif(not os.path.exists(os.path.join(REFERENCE, "GaussQuadratureTri.cc")), reason="reference not mounted")
@pytest.mark.parametrize("deg,name,n", [(2, "GQT2", 3), (3, "GQT3", 4), (4, "GQT4", 6), (5, "GQT5", 7), (6, "GQT6", 12),
                                        (7, "GQT7", 13), (8, "GQT8", 16), (9, "GQT9", 19), (10, "GQT10", 25)])
def test_tri_rule_matches_reference(deg, name, n):
    ref = _ref_tri_rule(name, n)
    ours = tri_rules()[deg]
    assert len(ours) == n
    # same point set (the reference prints 13-15 decimals)
    for p in ours:
        d = np.abs(ref[:, 0] - p[0]) + np.abs(ref[:, 1] - p[1])
        k = np.argmin(d)
        assert d[k] < 1e-12
        assert abs(ref[k, 2] - p[2]) < 1e-12
