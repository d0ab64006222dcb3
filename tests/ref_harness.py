"""Runs the reference's own DG classes (oracle/_ref/ref_driver, built from the
reference sources by oracle/Makefile.ref) on a mesh written as a Gmsh v1 file.
The reference's periodicity reflects x -> -x / y -> -y (DGCell.cc:557-577), which
equals translation on the origin-centred square [-1,1]^2 used here."""
import os
import subprocess
import tempfile

import numpy as np

from helpers import ROOT, unit_rand

DRIVER = os.path.join(ROOT, "oracle", "_ref", "ref_driver")


def available():
    return os.path.exists(DRIVER)


def square_mesh(n, jitter=0.1, seed=5, periodic=True, side_tags=(40000, 7, 40000, 7)):
    """Mesh of [-1,1]^2: lattice ids (unique per lattice vertex, for the reference)
    and wrapped topological ids (for the oracle / product)."""
    h = 2.0 / n
    N1 = n + 1
    coords = np.zeros((N1 * N1, 2))
    for j in range(N1):
        for i in range(N1):
            x, y = -1 + i * h, -1 + j * h
            if 0 < i < n and 0 < j < n and jitter > 0:
                vid = j * N1 + i
                x += (2.0 * unit_rand(seed ^ (2 * vid)) - 1.0) * jitter * h
                y += (2.0 * unit_rand(seed ^ (2 * vid + 1)) - 1.0) * jitter * h
            coords[j * N1 + i] = (x, y)

    def wrap(i, j):
        if periodic:
            i, j = i % n, j % n
        return j * N1 + i

    lat, top, xyz, bc = [], [], [], []
    for j in range(n):
        for i in range(n):
            for tri in (((i, j), (i + 1, j), (i + 1, j + 1)), ((i, j), (i + 1, j + 1), (i, j + 1))):
                lat.append([b * N1 + a for a, b in tri])
                top.append([wrap(a, b) for a, b in tri])
                xyz.append([coords[b * N1 + a] for a, b in tri])
                row = []
                for l in range(3):
                    (i0, j0), (i1, j1) = tri[l], tri[(l + 1) % 3]
                    side = 0 if (j0 == 0 and j1 == 0) else 1 if (i0 == n and i1 == n) else \
                        2 if (j0 == n and j1 == n) else 3 if (i0 == 0 and i1 == 0) else -1
                    row.append(0 if (side < 0 or periodic) else side_tags[side])
                bc.append(row)
    edges = []
    for i in range(n):
        edges.append(((i, 0), (i + 1, 0), 0))
        edges.append(((i, n), (i + 1, n), 2))
    for j in range(n):
        edges.append(((0, j), (0, j + 1), 3))
        edges.append(((n, j), (n, j + 1), 1))
    return dict(n=n, coords=coords, lat=np.array(lat), top=np.array(top, np.int32), xyz=np.array(xyz),
                bc=np.array(bc, np.int32), edges=edges, periodic=periodic, side_tags=side_tags)


def write_gmsh(mesh, path):
    n = mesh["n"]
    N1 = n + 1
    with open(path, "w") as f:
        f.write("$NOD\n%d\n" % len(mesh["coords"]))
        for k, (x, y) in enumerate(mesh["coords"]):
            f.write("%d %.17e %.17e 0\n" % (k + 1, x, y))
        f.write("$ENDNOD\n$ELM\n%d\n" % (len(mesh["edges"]) + len(mesh["lat"])))
        k = 1
        for (a, b, side) in mesh["edges"]:
            if mesh["periodic"]:
                tag = 510 if side in (1, 3) else 610
            else:
                tag = mesh["side_tags"][side]
            f.write("%d 1 %d 1 2 %d %d\n" % (k, tag, a[1] * N1 + a[0] + 1, b[1] * N1 + b[0] + 1))
            k += 1
        for t in mesh["lat"]:
            f.write("%d 2 100 1 3 %d %d %d\n" % (k, t[0] + 1, t[1] + 1, t[2] + 1))
            k += 1
        f.write("$ENDELM\n")


NFIELDS = {"scalar": 1, "euler": 4, "maxwell2d": 3, "burgers": 1}


def run(mesh, law, flux, order, op, U, dt=0.0, nsteps=1, vel=(1.0, 0.5), farfield=(0, 0, 0, 0), t0=0.0, far=None,
        stats=None):
    """Returns (values in the caller's element order, reference vertex order per element, resid list).
    vel: the scalar law's constant velocity, or "rotate" for ScalarPb's (2 pi y, -2 pi x).
    order: one order for every cell, or a per-element array (mixed orders; U is then
    the concatenation of the per-element blocks of Nf*Np(p_e) coefficients).
    far: "bump" for ScalarPb's initial condition as the scalar law's far field
    (op "project" returns the constructor's projection of it).
    stats: a dict that receives what the driver timed (REF_SWEEPS in os.environ)."""
    E = len(mesh["lat"])
    U = np.asarray(U, np.float64).ravel()
    mixed = not np.isscalar(order)
    orders = np.asarray(order, np.int64) if mixed else np.full(E, int(order))
    nf = NFIELDS[law]
    sizes = nf * (orders + 1) * (orders + 2) // 2
    offs = np.concatenate([[0], np.cumsum(sizes)])
    with tempfile.TemporaryDirectory() as d:
        mp, sp, op_ = os.path.join(d, "m.msh"), os.path.join(d, "s.txt"), os.path.join(d, "o.txt")
        write_gmsh(mesh, mp)
        with open(sp, "w") as f:
            for e in range(E):
                f.write(" ".join(str(v + 1) for v in mesh["lat"][e]) + (" %d" % orders[e] if mixed else "") + " " +
                        " ".join("%.17e" % x for x in U[offs[e]:offs[e + 1]]) + "\n")
        env = dict(os.environ, REF_VEL=vel if isinstance(vel, str) else "%r %r" % tuple(vel[:2]),
                   REF_FARFIELD=" ".join("%r" % x for x in farfield),
                   REF_T0="%r" % t0)
        if far:
            env["REF_FAR"] = far
        r = subprocess.run([DRIVER, mp, law, flux, "0" if mixed else str(order), op, sp, op_, "%r" % dt,
                            str(nsteps)], cwd=d, env=env, capture_output=True, text=True, timeout=600)
        if r.returncode != 0:
            raise RuntimeError("ref_driver failed: %s\n%s" % (r.stdout[-2000:], r.stderr[-2000:]))
        resid = [float(l.split()[1]) for l in r.stdout.splitlines() if l.startswith("resid ")]
        if stats is not None:
            for l in r.stdout.splitlines():
                if l.startswith("residual_seconds ") or l.startswith("step_seconds "):
                    stats[l.split()[0]] = float(l.split()[1])
        rows = [l.split() for l in open(op_).read().splitlines() if l.strip()]
    key = {tuple(sorted(mesh["lat"][e] + 1)): e for e in range(E)}
    out = np.zeros_like(U)
    order_v = np.zeros((E, 3), np.int64)
    for row in rows:
        v = [int(x) for x in row[:3]]
        e = key[tuple(sorted(v))]
        out[offs[e]:offs[e + 1]] = [float(x) for x in row[3:]]
        order_v[e] = v
    return out, order_v - 1, resid
