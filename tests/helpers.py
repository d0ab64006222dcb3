"""Test-side helpers: an independent numpy mesh generator (same specification as
dg_mesh_tri_structured, written separately so the product generator is checked
against it) and state builders."""
import math
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REFERENCE = "/root/reference"

M64 = (1 << 64) - 1


def splitmix64(x):
    z = (x + 0x9E3779B97F4A7C15) & M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def unit_rand(x):
    return (splitmix64(x & M64) >> 11) * (1.0 / 9007199254740992.0)


def mesh_tri(nx, ny, x0=0.0, y0=0.0, lx=1.0, ly=1.0, periodic_x=True, periodic_y=True, jitter=0.0, seed=0,
             side_tags=(0, 0, 0, 0)):
    hx, hy = lx / nx, ly / ny

    def lattice(i, j):
        x, y = x0 + i * hx, y0 + j * hy
        if 0 < i < nx and 0 < j < ny and jitter > 0:
            vid = j * (nx + 1) + i
            x += (2.0 * unit_rand(seed ^ (2 * vid)) - 1.0) * jitter * hx
            y += (2.0 * unit_rand(seed ^ (2 * vid + 1)) - 1.0) * jitter * hy
        return x, y

    def topo(i, j):
        if periodic_x and i == nx:
            i = 0
        if periodic_y and j == ny:
            j = 0
        return j * (nx + 1) + i

    def side(a, b):
        (i0, j0), (i1, j1) = a, b
        if j0 == 0 and j1 == 0:
            return 0
        if i0 == nx and i1 == nx:
            return 1
        if j0 == ny and j1 == ny:
            return 2
        if i0 == 0 and i1 == 0:
            return 3
        return -1

    E = 2 * nx * ny
    vids = np.zeros((E, 3), np.int32)
    xyz = np.zeros((E, 3, 2))
    bc = np.zeros((E, 3), np.int32)
    e = 0
    for j in range(ny):
        for i in range(nx):
            for tri in (((i, j), (i + 1, j), (i + 1, j + 1)), ((i, j), (i + 1, j + 1), (i, j + 1))):
                for v in range(3):
                    vids[e, v] = topo(*tri[v])
                    xyz[e, v] = lattice(*tri[v])
                for l in range(3):
                    sd = side(tri[l], tri[(l + 1) % 3])
                    per = periodic_y if sd in (0, 2) else periodic_x if sd in (1, 3) else True
                    bc[e, l] = side_tags[sd] if (sd >= 0 and not per) else 0
                e += 1
    return vids, xyz, bc


def np_tri(p):
    return (p + 1) * (p + 2) // 2


def random_modal_perturbation(U, ndof_per_elem, scale, seed):
    rng = np.random.default_rng(seed)
    return U + scale * rng.standard_normal(U.shape)


VORTEX = (5.0, 5.0, 5.0, 1.0, 1.0, 10.0)  # beta, xc, yc, u0, v0, L


def collapsed_rule(n):
    """Collapsed (Duffy) Gauss-Legendre rule on the reference triangle, n x n points,
    exact for polynomials of degree 2n-2; independent of the generated tables."""
    x, w = np.polynomial.legendre.leggauss(n)
    t = 0.5 * (x + 1)
    wt = 0.5 * w
    pts, wts = [], []
    for a, wa in zip(t, wt):
        for b, wb in zip(t, wt):
            u = a
            v = (1 - a) * b
            pts.append((u, v))
            wts.append(wa * wb * (1 - a))
    return np.array(pts), np.array(wts)


# far-field state off the vortex's free stream: where the two agree, the
# Lax-Friedrichs wave-speed maximum ties on the boundary faces and the discrete
# residual has a kink within reach of the finite-difference step
FF_OFF = (1.0, 0.8, 0.6, 3.0)


def roughen(U, p, seed=7):
    """the projected field plus a seeded perturbation of its three linear
    coefficients per element and field (0.5% of its scale): the face jumps and
    wave speeds of neighbouring traces then differ by far more than a
    finite-difference step, so no maximum or upwind switch ties within reach of
    it (as it would in an exactly uniform free stream)"""
    np_ = (p + 1) * (p + 2) // 2
    V = U.reshape(-1, np_).copy()
    V[:, :3] += 0.005 * np.abs(U).max() * np.random.default_rng(seed).standard_normal((V.shape[0], 3))
    return V.reshape(U.shape)


def direction(U, seed):
    """a random direction scaled per entry to the state (keeps the Euler states physical)"""
    rng = np.random.default_rng(seed)
    return rng.standard_normal(U.shape) * (0.05 * np.abs(U).max())


def write_msh(path, nodes, records):
    """A Gmsh version-1 file: nodes [(x, y, z)] numbered from 1, records
    [(type, physical tag, [node numbers])]."""
    with open(path, "w") as f:
        f.write("$NOD\n%d\n" % len(nodes))
        for i, (x, y, z) in enumerate(nodes):
            f.write("%d %.17g %.17g %.17g\n" % (i + 1, x, y, z))
        f.write("$ENDNOD\n$ELM\n%d\n" % len(records))
        for i, (typ, tag, nv) in enumerate(records):
            f.write("%d %d %d %d %d %s\n" % (i + 1, typ, tag, 1, len(nv), " ".join(str(v) for v in nv)))
        f.write("$ENDELM\n")


def lattice_msh(path, nx, ny, lx=1.0, ly=1.0, side_tags=(40000, 1, 40000, 1)):
    """The structured triangle mesh of dg_mesh_tri_structured (no jitter, cell
    (i,j) split by its (i,j)-(i+1,j+1) diagonal into (v00, v10, v11) and
    (v00, v11, v01)) as a Gmsh file with every lattice vertex its own node and the
    bottom, right, top, left sides tagged side_tags."""
    nid = lambda i, j: j * (nx + 1) + i + 1
    nodes = [(lx * i / nx, ly * j / ny, 0.0) for j in range(ny + 1) for i in range(nx + 1)]
    recs = []
    for j in range(ny):
        for i in range(nx):
            recs.append((2, 100, [nid(i, j), nid(i + 1, j), nid(i + 1, j + 1)]))
            recs.append((2, 100, [nid(i, j), nid(i + 1, j + 1), nid(i, j + 1)]))
    for i in range(nx):
        recs.append((1, side_tags[0], [nid(i, 0), nid(i + 1, 0)]))
        recs.append((1, side_tags[2], [nid(i + 1, ny), nid(i, ny)]))
    for j in range(ny):
        recs.append((1, side_tags[1], [nid(nx, j), nid(nx, j + 1)]))
        recs.append((1, side_tags[3], [nid(0, j + 1), nid(0, j)]))
    write_msh(path, nodes, recs)


def mesh_tri_graded(xs, ys, periodic=True):
    """A tensor grid of nodes xs x ys, each cell split into two triangles as
    mesh_tri splits them, periodic in x and y: cells of different sizes for
    local time stepping.  Returns (vids, xyz)."""
    xs, ys = np.asarray(xs, float), np.asarray(ys, float)
    nx, ny = len(xs) - 1, len(ys) - 1

    def topo(i, j):
        if periodic and i == nx:
            i = 0
        if periodic and j == ny:
            j = 0
        return j * (nx + 1) + i

    E = 2 * nx * ny
    vids = np.zeros((E, 3), np.int32)
    xyz = np.zeros((E, 3, 2))
    e = 0
    for j in range(ny):
        for i in range(nx):
            for tri in (((i, j), (i + 1, j), (i + 1, j + 1)), ((i, j), (i + 1, j + 1), (i, j + 1))):
                for v in range(3):
                    vids[e, v] = topo(*tri[v])
                    xyz[e, v] = (xs[tri[v][0]], ys[tri[v][1]])
                e += 1
    return vids, xyz


def graded_nodes(n, lo, hi, ratios):
    """n cells over [lo, hi] with widths proportional to ratios (cycled over the
    cells in blocks), e.g. ratios (4, 2, 1, 1, 2, 4) over 6 blocks"""
    nb = len(ratios)
    w = np.repeat(np.asarray(ratios, float), n // nb)
    w = np.concatenate([w, np.full(n - len(w), ratios[-1], float)])
    x = np.concatenate([[0.0], np.cumsum(w)])
    return lo + (hi - lo) * x / x[-1]


def refine_tri(vids, xyz, bc, mask):
    """Red refinement (each marked triangle into 4 through its edge midpoints),
    written independently of the library's refinement: corners P0, P1, P2 and
    midpoints M01, M12, M20 give the children (P0, M01, M20), (M01, P1, M12),
    (M20, M12, P2), (M12, M20, M01), in place of the parent, in order.  A
    midpoint's topological id is new per edge (shared by both sides' images);
    its coordinates come from each element's own corners.  Halves of a parent's
    face keep its tag.  Unmarked neighbours keep their faces (hanging nodes).
    Returns (vids, xyz, bc, parent) with parent[i] the input element of output i."""
    vids = np.asarray(vids)
    xyz = np.asarray(xyz, float)
    bc = np.zeros_like(vids) if bc is None else np.asarray(bc)
    nxt = int(vids.max()) + 1
    mids = {}

    def mid(a, b):
        nonlocal nxt
        k = (min(a, b), max(a, b))
        if k not in mids:
            mids[k] = nxt
            nxt += 1
        return mids[k]

    ov, ox, ob, par = [], [], [], []
    for e in range(len(vids)):
        if not mask[e]:
            ov.append(vids[e]), ox.append(xyz[e]), ob.append(bc[e]), par.append(e)
            continue
        P = list(vids[e])
        X = list(xyz[e])
        m01, m12, m20 = mid(P[0], P[1]), mid(P[1], P[2]), mid(P[2], P[0])
        X01, X12, X20 = 0.5 * (X[0] + X[1]), 0.5 * (X[1] + X[2]), 0.5 * (X[2] + X[0])
        t = bc[e]
        kids = [((P[0], m01, m20), (X[0], X01, X20), (t[0], 0, t[2])),
                ((m01, P[1], m12), (X01, X[1], X12), (t[0], t[1], 0)),
                ((m20, m12, P[2]), (X20, X12, X[2]), (0, t[1], t[2])),
                ((m12, m20, m01), (X12, X20, X01), (0, 0, 0))]
        for v, x, b in kids:
            ov.append(v), ox.append(np.array(x)), ob.append(b), par.append(e)
    return (np.array(ov, np.int32), np.array(ox), np.array(ob, np.int32), np.array(par, np.int32))
