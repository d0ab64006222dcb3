#!/usr/bin/env python3
"""Generate csrc/dg_rules_data.h: quadrature rules and orthonormal modal-basis
coefficients, computed from their defining equations in 60-digit arithmetic and
rounded once to double.

What is generated, and which reference routine each table stands in for:

* Triangle rules of degree 1..10 (reference selection `GaussQuadratureTri.cc:138-167`:
  degree d -> rule GQT[d] with 1,1,3,4,6,7,12,13,16,19,25 points).  Each rule is the
  symmetric rule with the published orbit structure (Strang-Fix degree 3, Radon
  degree 5, Dunavant 1985 for the others).  Points and weights are solved here by
  Gauss-Newton on the moment equations sum_q w_q u^i v^j = 2 i! j! / (i+j+2)!,
  started from 4-digit orbit parameters; weights sum to 1 (the caller scales by
  the reference-element measure 1/2, as `Integrator.cc:69` does).
* Tetrahedron rules of degree 1..6 (`GaussQuadratureTet.cc:349-380`: 1,1,4,5,14,17,24
  points; the 24-point degree-6 rule is Keast's), same procedure, weights sum to 1
  (caller scales by 1/6, `Integrator.cc:97`).
* Gauss-Legendre rules on [-1,1] with 1..12 points (`GaussLegendre1D.cc:163-274`;
  degree -> points map `getNGQ1DPts`, `GaussLegendre1D.cc:271-274`).
* Gram-Schmidt orthonormalisation of the Pascal monomials on the reference
  triangle (monomial order `ShapeFunctionList.cc:311-349`) and tetrahedron
  (`ShapeFunctionList.cc:666-706`), as lower-triangular matrices C with
  phi_i = sum_{j<=i} C[i][j] m_j and C[i][i] > 0 -- the same construction as the
  reference's offline generator (`DGOrthogonalize.cc:17-163`), done exactly
  (Cholesky factor of the monomial Gram matrix, inverted) instead of being read
  from printed tables.

Run:  python gen_rules.py  (writes csrc/dg_rules_data.h next to this file).
"""
import os
import sys

import mpmath as mp

mp.mp.dps = 60
TOL = mp.mpf(10) ** -45


# ----------------------------------------------------------------------------
# moments
# ----------------------------------------------------------------------------
def tri_moment(i, j):
    # integral over the reference triangle of u^i v^j, times 2 (weights sum to 1)
    return 2 * mp.factorial(i) * mp.factorial(j) / mp.factorial(i + j + 2)


def tet_moment(i, j, k):
    return 6 * mp.factorial(i) * mp.factorial(j) * mp.factorial(k) / mp.factorial(i + j + k + 3)


# ----------------------------------------------------------------------------
# orbit expansions (barycentric -> (u, v[, w]) points)
# ----------------------------------------------------------------------------
def tri_orbit(kind, params):
    if kind == "S3":
        t = mp.mpf(1) / 3
        return [(t, t)]
    if kind == "S21":
        a, = params
        b = 1 - 2 * a
        return [(a, a), (a, b), (b, a)]
    if kind == "S111":
        a, b = params
        c = 1 - a - b
        return [(a, b), (b, a), (a, c), (c, a), (b, c), (c, b)]
    raise ValueError(kind)


def tet_orbit(kind, params):
    if kind == "S4":
        q = mp.mpf(1) / 4
        return [(q, q, q)]
    if kind == "S31":
        a, = params
        b = 1 - 3 * a
        return [(a, a, a), (a, a, b), (a, b, a), (b, a, a)]
    if kind == "S22":
        a, = params
        b = mp.mpf(1) / 2 - a
        return [(a, a, b), (a, b, a), (b, a, a), (b, b, a), (b, a, b), (a, b, b)]
    if kind == "S211":
        a, b = params
        c = 1 - 2 * a - b
        pts = []
        # barycentric permutations of (a, a, b, c); (u, v, w) are the first three
        seen = set()
        import itertools
        for perm in itertools.permutations((0, 0, 1, 2)):
            if perm in seen:
                continue
            seen.add(perm)
            val = [a, b, c]
            pts.append(tuple(val[perm[i]] for i in range(3)))
        assert len(pts) == 12
        return pts
    raise ValueError(kind)


# ----------------------------------------------------------------------------
# Gauss-Newton on the moment equations
# ----------------------------------------------------------------------------
def solve_rule(structure, degree, dim):
    """structure: list of (kind, [start params], start weight or None)."""
    if dim == 2:
        orbit = tri_orbit
        monos = [(i, n - i) for n in range(degree + 1) for i in range(n + 1)]
        moment = lambda m: tri_moment(*m)
        ev = lambda p, m: p[0] ** m[0] * p[1] ** m[1]
    else:
        orbit = tet_orbit
        monos = [(i, j, n - i - j) for n in range(degree + 1) for i in range(n + 1) for j in range(n + 1 - i)]
        moment = lambda m: tet_moment(*m)
        ev = lambda p, m: p[0] ** m[0] * p[1] ** m[1] * p[2] ** m[2]

    kinds = [s[0] for s in structure]
    nparams = [len(s[1]) for s in structure]
    x = []
    for s in structure:
        x += [mp.mpf(v) for v in s[1]]
    norb = len(structure)

    def unpack(xv):
        out, o = [], 0
        for n in nparams:
            out.append(xv[o:o + n])
            o += n
        return out, xv[o:]

    def orbit_sums(params):
        # S[o][m] = sum over the points of orbit o of monomial m
        return [[mp.fsum(ev(p, m) for p in orbit(kinds[o], params[o])) for m in monos] for o in range(norb)]

    # initial weights: linear least squares with the starting points
    params, _ = unpack(x)
    S = orbit_sums(params)
    A = mp.matrix(len(monos), norb)
    rhs = mp.matrix(len(monos), 1)
    for r, m in enumerate(monos):
        for o in range(norb):
            A[r, o] = S[o][r]
        rhs[r] = moment(m)
    w0 = mp.lu_solve(A.T * A, A.T * rhs)
    x = x + [w0[o] for o in range(norb)]

    def residual(xv):
        params, w = unpack(xv)
        S = orbit_sums(params)
        return mp.matrix([mp.fsum(w[o] * S[o][r] for o in range(norb)) - moment(m) for r, m in enumerate(monos)])

    h = mp.mpf(10) ** -25
    for it in range(60):
        r = residual(x)
        nr = mp.norm(r)
        if nr < TOL:
            break
        J = mp.matrix(len(monos), len(x))
        for c in range(len(x)):
            xp = list(x)
            xm = list(x)
            xp[c] += h
            xm[c] -= h
            d = (residual(xp) - residual(xm)) / (2 * h)
            for rr in range(len(monos)):
                J[rr, c] = d[rr]
        dx = mp.lu_solve(J.T * J, -(J.T * r))
        x = [x[c] + dx[c] for c in range(len(x))]
    r = residual(x)
    if mp.norm(r) > mp.mpf(10) ** -40:
        raise RuntimeError("rule of degree %d did not converge: |r| = %s" % (degree, mp.nstr(mp.norm(r), 5)))
    params, w = unpack(x)
    pts = []
    for o in range(norb):
        for p in orbit(kinds[o], params[o]):
            pts.append(tuple(p) + (w[o],))
    for p in pts:
        for c in list(p[:dim]) + [1 - sum(p[:dim])]:
            if not (-TOL < c < 1 + TOL):
                raise RuntimeError("point outside the element in degree %d" % degree)
    return pts


# orbit structures with rough (4-digit) starting parameters
TRI_RULES = {
    1: [("S3", [], None)],
    2: [("S21", ["0.1667"], None)],
    3: [("S3", [], None), ("S21", ["0.2001"], None)],
    4: [("S21", ["0.0916"], None), ("S21", ["0.4459"], None)],
    5: [("S3", [], None), ("S21", ["0.1013"], None), ("S21", ["0.4701"], None)],
    6: [("S21", ["0.0631"], None), ("S21", ["0.2493"], None), ("S111", ["0.0531", "0.3104"], None)],
    7: [("S3", [], None), ("S21", ["0.2603"], None), ("S21", ["0.0651"], None), ("S111", ["0.0487", "0.3129"], None)],
    8: [("S3", [], None), ("S21", ["0.4593"], None), ("S21", ["0.1706"], None), ("S21", ["0.0505"], None),
        ("S111", ["0.0084", "0.2631"], None)],
    9: [("S3", [], None), ("S21", ["0.4897"], None), ("S21", ["0.4371"], None), ("S21", ["0.1882"], None),
        ("S21", ["0.0447"], None), ("S111", ["0.0368", "0.2220"], None)],
    10: [("S3", [], None), ("S21", ["0.4856"], None), ("S21", ["0.1095"], None),
         ("S111", ["0.1417", "0.3079"], None), ("S111", ["0.0250", "0.2467"], None),
         ("S111", ["0.0095", "0.0668"], None)],
}

TET_RULES = {
    1: [("S4", [], None)],
    2: [("S31", ["0.1382"], None)],
    3: [("S4", [], None), ("S31", ["0.1667"], None)],
    4: [("S22", ["0.0"], None), ("S31", ["0.1005"], None), ("S31", ["0.3143"], None)],
    6: [("S31", ["0.2146"], None), ("S31", ["0.0407"], None), ("S31", ["0.3223"], None),
        ("S211", ["0.0637", "0.2697"], None)],
}


def gauss_legendre(n):
    pts = []
    for i in range(1, n + 1):
        x = mp.cos(mp.pi * (i - mp.mpf(1) / 4) / (n + mp.mpf(1) / 2))
        for _ in range(100):
            p0, p1 = mp.mpf(1), x
            for k in range(2, n + 1):
                p0, p1 = p1, ((2 * k - 1) * x * p1 - (k - 1) * p0) / k
            if n == 1:
                pn, pnm1 = x, mp.mpf(1)
            else:
                pn, pnm1 = p1, p0
            dp = n * (x * pn - pnm1) / (x * x - 1)
            dx = pn / dp
            x -= dx
            if abs(dx) < mp.mpf(10) ** -55:
                break
        p0, p1 = mp.mpf(1), x
        for k in range(2, n + 1):
            p0, p1 = p1, ((2 * k - 1) * x * p1 - (k - 1) * p0) / k
        pn, pnm1 = (x, mp.mpf(1)) if n == 1 else (p1, p0)
        dp = n * (x * pn - pnm1) / (x * x - 1)
        w = 2 / ((1 - x * x) * dp * dp)
        pts.append((x, w))
    return sorted(pts)


# Pascal monomial exponents in the reference order
def tri_monomials(nmax):
    out = []
    n = 0
    while len(out) < nmax:
        for i in range(n + 1):
            out.append((n - i, i))
        n += 1
    return out[:nmax]


TET_MONOMIALS = [
    (0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1),
    (2, 0, 0), (0, 2, 0), (0, 0, 2), (1, 1, 0), (1, 0, 1), (0, 1, 1),
    (3, 0, 0), (0, 3, 0), (0, 0, 3), (2, 1, 0), (2, 0, 1), (1, 2, 0), (0, 2, 1), (1, 0, 2), (0, 1, 2), (1, 1, 1),
    (4, 0, 0), (0, 4, 0), (0, 0, 4), (3, 1, 0), (3, 0, 1), (1, 3, 0), (0, 3, 1), (1, 0, 3), (0, 1, 3),
    (2, 2, 0), (2, 0, 2), (0, 2, 2), (2, 1, 1), (1, 2, 1), (1, 1, 2),
]


def gram_schmidt(monos, integ):
    n = len(monos)
    G = mp.matrix(n, n)
    for i in range(n):
        for j in range(n):
            G[i, j] = integ(tuple(a + b for a, b in zip(monos[i], monos[j])))
    L = mp.cholesky(G)
    C = mp.inverse(L)
    return C


def fmt(x):
    return "%.17e" % float(x)


def main():
    out = []
    out.append("/* Generated by gen_rules.py -- do not edit.  Quadrature rules and orthonormal")
    out.append(" * modal-basis coefficients solved in 60-digit arithmetic, rounded to double. */")
    out.append("#pragma once")
    out.append("")
    # triangle rules
    offs, npts, rows = [], [], []
    for d in range(1, 11):
        tri_cache[d] = solve_rule(TRI_RULES[d], d, 2)
    for d in range(0, 11):
        pts = tri_cache[max(d, 1)]
        offs.append(len(rows))
        npts.append(len(pts))
        rows += pts
    out.append("#define DG_TRI_RULE_MAXDEG 10")
    out.append("static const int dg_tri_rule_off[11] = {%s};" % ", ".join(map(str, offs)))
    out.append("static const int dg_tri_rule_npt[11] = {%s};" % ", ".join(map(str, npts)))
    out.append("/* {u, v, weight}; weights sum to 1 */")
    out.append("static const double dg_tri_rule_pts[%d][3] = {" % len(rows))
    for r in rows:
        out.append("  {%s, %s, %s}," % (fmt(r[0]), fmt(r[1]), fmt(r[2])))
    out.append("};")
    out.append("")
    # tet rules
    offs, npts, rows = [], [], []
    for d in TET_RULES:
        tet_cache[d] = solve_rule(TET_RULES[d], d, 3)
    for d in range(0, 7):
        # degree 5 is served by the degree-6 rule (no volume rule of odd degree is used)
        pts = tet_cache[{0: 1, 5: 6}.get(d, d)]
        offs.append(len(rows))
        npts.append(len(pts))
        rows += pts
    out.append("#define DG_TET_RULE_MAXDEG 6")
    out.append("static const int dg_tet_rule_off[7] = {%s};" % ", ".join(map(str, offs)))
    out.append("static const int dg_tet_rule_npt[7] = {%s};" % ", ".join(map(str, npts)))
    out.append("/* {u, v, w, weight}; weights sum to 1 */")
    out.append("static const double dg_tet_rule_pts[%d][4] = {" % len(rows))
    for r in rows:
        out.append("  {%s, %s, %s, %s}," % (fmt(r[0]), fmt(r[1]), fmt(r[2]), fmt(r[3])))
    out.append("};")
    out.append("")
    # Gauss-Legendre
    offs, rows = [], []
    for n in range(1, 13):
        offs.append(len(rows))
        rows += gauss_legendre(n)
    out.append("#define DG_GL_MAXPTS 12")
    out.append("/* rule with n points starts at dg_gl_off[n-1]; {x in [-1,1], weight}; weights sum to 2 */")
    out.append("static const int dg_gl_off[12] = {%s};" % ", ".join(map(str, offs)))
    out.append("static const double dg_gl_pts[%d][2] = {" % len(rows))
    for r in rows:
        out.append("  {%s, %s}," % (fmt(r[0]), fmt(r[1])))
    out.append("};")
    out.append("")
    # Gram-Schmidt
    ntri = 28  # up to P6
    tm = tri_monomials(ntri)
    C = gram_schmidt(tm, lambda e: tri_moment(*e) / 2)
    out.append("#define DG_TRI_MAXMODES %d" % ntri)
    out.append("static const int dg_tri_mono[%d][2] = {%s};" % (ntri, ", ".join("{%d,%d}" % m for m in tm)))
    out.append("/* phi_i(u,v) = sum_{j<=i} C[i][j] u^a_j v^b_j, orthonormal on the reference triangle */")
    out.append("static const double dg_tri_gs[%d][%d] = {" % (ntri, ntri))
    for i in range(ntri):
        out.append("  {" + ", ".join(fmt(C[i, j]) if j <= i else "0.0" for j in range(ntri)) + "},")
    out.append("};")
    out.append("")
    ntet = len(TET_MONOMIALS)
    C = gram_schmidt(TET_MONOMIALS, lambda e: tet_moment(*e) / 6)
    out.append("#define DG_TET_MAXMODES %d" % ntet)
    out.append("static const int dg_tet_mono[%d][3] = {%s};" % (ntet, ", ".join("{%d,%d,%d}" % m for m in TET_MONOMIALS)))
    out.append("/* phi_i(u,v,w) = sum_{j<=i} C[i][j] u^a_j v^b_j w^c_j, orthonormal on the reference tetrahedron */")
    out.append("static const double dg_tet_gs[%d][%d] = {" % (ntet, ntet))
    for i in range(ntet):
        out.append("  {" + ", ".join(fmt(C[i, j]) if j <= i else "0.0" for j in range(ntet)) + "},")
    out.append("};")
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "csrc", "dg_rules_data.h")
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")
    print("wrote", path)


tri_cache = {}
tet_cache = {}

if __name__ == "__main__":
    sys.exit(main())
