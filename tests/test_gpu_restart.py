"""GPU: checkpoint / restart (dg_write_restart, dg_read_restart: DGAnalysis::write
and ::read, DGPref.cc:712-768) and state export (dg_sample_lattice,
dg_export_pos: DGAnalysis::exportGmsh, DGSensor.cc:337-525, 800-).

- a restart file holds the reference's record layout (t, iteration, "dim E F",
  per cell the entity type, vertex ids, order and Nf*Np coefficients field by
  field, then one record per face) and reads back to the same doubles;
- a run stopped, written, read into a fresh context and continued ends on the
  same state, bit for bit, as the run without the stop;
- a mixed-order state read into a uniform-order context re-buckets it to the
  file's orders; records in another element order are matched by vertex ids;
  mismatches (element count, vertex order, truncation, orders of a form that
  cannot change them) fail with a message;
- the mesh written next to the data (DGAnalysis::write(mesh, data)) reads back
  through the Gmsh reader into a context that takes the data file;
- the lattice sampler matches the oracle's basis applied to the state (1e-13)
  and returns a constant for a constant state (the mean mode only);
- the .pos views carry the sampled fields at the corners of the plot
  sub-simplices: density, pressure, velocity and Mach from the sampled
  conserved values, E*nsub^2 triangles or E*8 tetrahedra per view."""
import ctypes
import os
import re
import sys

import numpy as np
import pytest

from helpers import VORTEX, mesh_tri, np_tri
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def announce(name):
    print("\n[gpu case] %s" % name, flush=True)
    sys.stdout.flush()


def vortex(p=3, n=8, precision=0, orders=None):
    import dgpar
    vids, xyz, bc = mesh_tri(n, n, lx=10.0, ly=10.0, jitter=0.1, seed=3)
    m = dgpar.Mesh(2, vids, xyz, bc, order=orders)
    g = dgpar.DGAnalysis(m, dgpar.EulerLaw(dgpar.FLUX_LF, farfield=(1.0, 1.0, 1.0, 3.5)), p, precision=precision)
    g.project(dgpar.FIELD_VORTEX, VORTEX)
    return m, g


def tokens(path):
    with open(path) as f:
        return f.read().split()


def test_restart_file_layout_and_round_trip(tmp_path):
    import dgpar
    announce("restart layout, P3 vortex")
    m, g = vortex()
    dt = g.computeDt(0.1)
    g.advanceInTime(0.0, dt, scheme=dgpar.RK4, nsteps=2)
    U = g.get_state()
    p = tmp_path / "run.dat"
    g.write_restart(p, 2 * dt, 2)
    tk = tokens(p)
    E, F = m.n_elems, g.faces().shape[0]
    assert float(tk[0]) == 2 * dt and tk[1:5] == ["2", "2", str(E), str(F)]
    # first cell: type TRI = 2, its vertex ids as node numbers (id + 1), order 3,
    # 4 x 10 coefficients
    assert tk[5] == "2" and [int(v) for v in tk[6:9]] == (m.vids[0] + 1).tolist() and tk[9] == "3"
    assert np.array([float(v) for v in tk[10:50]]).tolist() == U[:40].tolist()
    # after the cells, F edge records (type 1 and two vertex ids)
    cells = 5 + E * (1 + 3 + 1 + 40)
    assert len(tk) == cells + 3 * F and all(tk[cells + 3 * i] == "1" for i in range(F))
    _, h = vortex()
    t, it = h.read_restart(p)
    assert t == 2 * dt and it == 2
    assert (h.get_state() == U).all()


def test_restart_continues_bit_for_bit(tmp_path):
    import dgpar
    announce("restart continuation, RK4 and FP32 form")
    for prec in (0, 1):
        _, a = vortex(precision=prec)
        dt = a.computeDt(0.1)
        a.advanceInTime(0.0, dt, scheme=dgpar.RK4, nsteps=3)
        a.write_restart(tmp_path / "mid.dat", 3 * dt, 3)
        a.advanceInTime(3 * dt, dt, scheme=dgpar.RK4, nsteps=3)
        _, b = vortex(precision=prec)
        t, it = b.read_restart(tmp_path / "mid.dat")
        b.advanceInTime(t, dt, scheme=dgpar.RK4, nsteps=3)
        assert (a.get_state() == b.get_state()).all()


def test_restart_takes_the_files_orders(tmp_path):
    import dgpar
    announce("restart of mixed orders into a uniform context")
    vids, xyz, bc = mesh_tri(8, 8, lx=10.0, ly=10.0, jitter=0.1, seed=3)
    orders = np.random.default_rng(2).integers(1, 5, vids.shape[0]).astype(np.int32)
    _, a = vortex(p=2, orders=orders)
    U = a.get_state()
    a.write_restart(tmp_path / "mix.dat", 0.5, 7)
    _, b = vortex(p=2)
    assert (b.orders() == 2).all()
    assert b.read_restart(tmp_path / "mix.dat") == (0.5, 7)
    assert (b.orders() == orders).all() and b.ndof == U.size
    assert (b.get_state() == U).all()
    # the re-bucketed context computes the same residual as one built on these orders
    assert np.abs(b.residual() - a.residual()).max() <= 1e-12 * np.abs(a.residual()).max()
    # the FP32 form has one order: a file with other orders is refused
    _, c = vortex(p=2, precision=1)
    with pytest.raises(dgpar.DGError, match="FP64 form"):
        c.read_restart(tmp_path / "mix.dat")


def test_restart_matches_cells_by_vertex_ids(tmp_path):
    import dgpar
    announce("restart with permuted element order")
    m, a = vortex(p=2)
    U = a.get_state().reshape(m.n_elems, -1)
    a.write_restart(tmp_path / "a.dat")
    perm = np.random.default_rng(5).permutation(m.n_elems)
    mp = dgpar.Mesh(2, m.vids[perm], m.xyz[perm], m.bc[perm])
    b = dgpar.DGAnalysis(mp, dgpar.EulerLaw(dgpar.FLUX_LF, farfield=(1.0, 1.0, 1.0, 3.5)), 2)
    b.read_restart(tmp_path / "a.dat")
    assert (b.get_state().reshape(m.n_elems, -1) == U[perm]).all()
    # an element whose vertices come in another (rotated) order holds other
    # coefficients: refused
    vr, xr = m.vids.copy(), m.xyz.copy()
    vr[0], xr[0] = np.roll(vr[0], 1), np.roll(xr[0], 1, axis=0)
    c = dgpar.DGAnalysis(dgpar.Mesh(2, vr, xr, m.bc), dgpar.EulerLaw(dgpar.FLUX_LF, farfield=(1.0, 1.0, 1.0, 3.5)), 2)
    with pytest.raises(dgpar.DGError, match="vertex order"):
        c.read_restart(tmp_path / "a.dat")
    # another mesh size, a truncated file
    _, d = vortex(p=2, n=6)
    with pytest.raises(dgpar.DGError, match="elements"):
        d.read_restart(tmp_path / "a.dat")
    txt = open(tmp_path / "a.dat").read()
    open(tmp_path / "t.dat", "w").write(txt[: len(txt) // 3])
    with pytest.raises(dgpar.DGError, match="truncated|bad"):
        b.read_restart(tmp_path / "t.dat")
    with pytest.raises(dgpar.DGError, match="cannot open"):
        b.read_restart(tmp_path / "missing.dat")


def test_reads_a_restart_the_reference_wrote():
    """tests/golden/ref_restart_tri.dat: DGAnalysis::write of the reference build
    (P2 Euler, a 6x4 lattice mesh file with walls and far field, node numbers as
    vertex ids) into a context built from that mesh file"""
    import dgpar
    announce("restart written by the reference")
    gold = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    m, info = dgpar.mesh_read_gmsh(os.path.join(gold, "ref_restart_tri.msh"))
    g = dgpar.DGAnalysis(m, dgpar.EulerLaw(dgpar.FLUX_ROE, farfield=(1.0, 0.2, -0.1, 2.6)), 2)
    t, it = g.read_restart(os.path.join(gold, "ref_restart_tri.dat"))
    assert t == 0.0 and it == 0
    # the records parsed here: (type 2, three node numbers, order 2, 24 coefficients)
    tk = tokens(os.path.join(gold, "ref_restart_tri.dat"))
    E = int(tk[3])
    assert E == m.n_elems
    want = {}
    i = 5
    for _ in range(E):
        assert tk[i] == "2" and tk[i + 4] == "2"
        want[tuple(int(v) for v in tk[i + 1:i + 4])] = [float(x) for x in tk[i + 5:i + 29]]
        i += 29
    U = g.get_state().reshape(E, 24)
    for e in range(E):
        assert U[e].tolist() == want[tuple((m.vids[e] + 1).tolist())]


def test_write_mesh_and_data_then_restart_from_files(tmp_path):
    import dgpar
    announce("DGAnalysis.write(mesh, data) on a periodic vortex, read back")
    m, a = vortex(p=3)
    dt = a.computeDt(0.1)
    a.advanceInTime(0.0, dt, scheme=dgpar.RK4, nsteps=2)
    a.write(tmp_path / "v.msh", tmp_path / "v.dat", 2 * dt, 2)
    r, _ = dgpar.mesh_read_gmsh(tmp_path / "v.msh")
    b = dgpar.DGAnalysis(r, dgpar.EulerLaw(dgpar.FLUX_LF, farfield=(1.0, 1.0, 1.0, 3.5)), 3)
    t, it = b.read_restart(tmp_path / "v.dat")
    a.advanceInTime(2 * dt, dt, scheme=dgpar.RK4, nsteps=2)
    b.advanceInTime(t, dt, scheme=dgpar.RK4, nsteps=2)
    assert (a.get_state() == b.get_state()).all()


def test_tetrahedra_restart(tmp_path):
    import dgpar
    announce("Maxwell tetrahedra restart")
    k = dgpar.mesh_tet_structured(3, 3, 3)
    g = dgpar.DGAnalysis(k, dgpar.MaxwellLaw(), 3)
    g.project(dgpar.FIELD_CAVITY, ())
    dt = g.computeDt(0.1)
    g.advanceInTime(0.0, dt, scheme=dgpar.RK4, nsteps=2)
    g.write_restart(tmp_path / "m.dat", 2 * dt, 2)
    tk = tokens(tmp_path / "m.dat")
    assert tk[2:4] == ["3", str(k.n_elems)] and tk[5] == "4" and tk[10] == "3"
    h = dgpar.DGAnalysis(k, dgpar.MaxwellLaw(), 3)
    h.read_restart(tmp_path / "m.dat")
    assert (h.get_state() == g.get_state()).all()
    h2 = dgpar.DGAnalysis(k, dgpar.MaxwellLaw(), 2)
    with pytest.raises(dgpar.DGError, match="one order"):
        h2.read_restart(tmp_path / "m.dat")


def lattice(dim, nsub):
    pts = []
    for kk in range(nsub + 1 if dim == 3 else 1):
        for j in range(nsub + 1 - kk):
            for i in range(nsub + 1 - kk - j):
                pts.append((i / nsub, j / nsub, kk / nsub))
    return np.array(pts)


def oracle_samples(U, orders, nf, dim, nsub):
    pts = lattice(dim, nsub)
    out, i = [], 0
    for p in orders:
        n = np_tri(p) if dim == 2 else (p + 1) * (p + 2) * (p + 3) // 6
        B = np.array([po.basis(dim, int(p), x)[0] for x in pts])  # [points][modes]
        out.append(U[i:i + nf * n].reshape(nf, n) @ B.T)
        i += nf * n
    return np.array(out)


@pytest.mark.parametrize("nsub", [1, 3, 8])
def test_lattice_sampler_matches_the_oracle(nsub):
    import dgpar
    announce("lattice sampler nsub %d" % nsub)
    vids, xyz, bc = mesh_tri(8, 8, lx=10.0, ly=10.0, jitter=0.1, seed=3)
    orders = np.random.default_rng(4).integers(1, 5, vids.shape[0]).astype(np.int32)
    _, g = vortex(p=2, orders=orders)
    U = g.get_state()
    S = g.sample_lattice(nsub)
    ref = oracle_samples(U, orders, 4, 2, nsub)
    assert np.abs(S - ref).max() <= 1e-13 * np.abs(ref).max()
    # the mean mode alone is a constant: phi_0 = sqrt(2) on the reference triangle
    V = np.zeros_like(U)
    i = 0
    for p in orders:
        V[i:i + 4 * np_tri(p):np_tri(p)] = np.array([1.0, 0.5, -0.25, 2.0]) / np.sqrt(2.0)
        i += 4 * np_tri(p)
    g.set_state(V)
    S = g.sample_lattice(nsub)
    assert np.abs(S - np.array([1.0, 0.5, -0.25, 2.0])[None, :, None]).max() <= 1e-15


def test_lattice_sampler_tetrahedra():
    import dgpar
    announce("lattice sampler tetrahedra P3")
    k = dgpar.mesh_tet_structured(3, 3, 3)
    g = dgpar.DGAnalysis(k, dgpar.MaxwellLaw(), 3)
    U = np.random.default_rng(8).standard_normal(g.ndof)
    g.set_state(U)
    S = g.sample_lattice(2)
    ref = oracle_samples(U, [3] * k.n_elems, 6, 3, 2)
    assert np.abs(S - ref).max() <= 1e-13 * np.abs(ref).max()
    with pytest.raises(dgpar.DGError, match="nsub"):
        g.sample_lattice(9)


NUM = r"[-+0-9.eE]+|nan|inf"


def pos_records(path):
    with open(path) as f:
        lines = f.read().splitlines()
    assert lines[0] == 'View "Exported field" {' and lines[-1] == "};"
    recs = []
    for ln in lines[1:-1]:
        m = re.match(r"^(ST|VT|SS|VS) \((.*)\) \{ (.*)\};$", ln)
        assert m, ln
        recs.append((m.group(1), np.array([float(v) for v in m.group(2).split(",")]),
                     np.array([float(v) for v in m.group(3).split(",")])))
    return recs


def test_pos_views_euler(tmp_path):
    import dgpar
    announce("pos export Euler P2")
    m, g = vortex(p=2, n=6)
    prefix = str(tmp_path / "vortex")
    g.export_pos(prefix, iteration=4)
    names = ["velocity", "pressure", "mach", "density", "error", "entropy", "limit", "order"]
    for nm in names:
        assert os.path.exists("%s-sample4%s.pos" % (prefix, nm)), nm
    S = g.sample_lattice(2)
    pts = lattice(2, 2)
    # the first element's sub-triangles: (0,1,3) (1,4,3) (1,2,4) (3,4,5) in lattice indices
    subs = [(0, 1, 3), (1, 4, 3), (1, 2, 4), (3, 4, 5)]
    X = m.xyz[0]
    rho = pos_records("%s-sample4density.pos" % prefix)
    assert len(rho) == m.n_elems * 4 and all(r[0] == "ST" for r in rho)
    for s, (tag, xs, vs) in zip(subs, rho[:4]):
        corners = np.array([X[0] + (X[1] - X[0]) * pts[q][0] + (X[2] - X[0]) * pts[q][1] for q in s])
        assert np.abs(xs.reshape(3, 3)[:, :2] - corners).max() <= 1e-9
        assert np.abs(vs - S[0, 0, list(s)]).max() <= 1e-9 * np.abs(vs).max()
    u = S[0][:, [0, 1, 3]]  # conserved values at the first sub-triangle's corners
    p = 0.4 * (u[3] - 0.5 * (u[1] ** 2 + u[2] ** 2) / u[0])
    pr = pos_records("%s-sample4pressure.pos" % prefix)[0][2]
    assert np.abs(pr - p).max() <= 1e-8 * p.max()
    ma = pos_records("%s-sample4mach.pos" % prefix)[0][2]
    vel = np.sqrt(u[1] ** 2 + u[2] ** 2) / u[0]
    assert np.abs(ma - vel / np.sqrt(1.4 * p / u[0])).max() <= 1e-8
    v = pos_records("%s-sample4velocity.pos" % prefix)[0]
    assert v[0] == "VT" and np.abs(v[2].reshape(3, 3) - np.c_[u[1] / u[0], u[2] / u[0], np.zeros(3)]).max() <= 1e-8
    od = pos_records("%s-sample4order.pos" % prefix)
    assert all((r[2] == 2).all() for r in od)


def test_pos_views_scalar_and_tetrahedra(tmp_path):
    import dgpar
    announce("pos export scalar with exact field, Maxwell tetrahedra")
    vids, xyz, bc = mesh_tri(6, 6)
    g = dgpar.DGAnalysis(dgpar.Mesh(2, vids, xyz, bc), dgpar.ScalarLaw((1.0, 0.5)), 3)
    g.project(dgpar.FIELD_GAUSSIAN, ())
    g.export_pos(str(tmp_path / "s"), 0, nsub=2, t=0.0, exact_field=dgpar.FIELD_GAUSSIAN)
    err = pos_records(str(tmp_path / "s-sample0error.pos"))
    ex = pos_records(str(tmp_path / "s-sample0exact.pos"))
    fn = pos_records(str(tmp_path / "s-sample0function.pos"))
    assert len(err) == 72 * 4
    d = np.array([e[2] - (x[2] - f[2]) for e, x, f in zip(err, ex, fn)])
    assert np.abs(d).max() <= 1e-8
    k = dgpar.mesh_tet_structured(2, 2, 2)
    h = dgpar.DGAnalysis(k, dgpar.MaxwellLaw(), 3)
    h.project(dgpar.FIELD_CAVITY, ())
    h.export_pos(str(tmp_path / "m"), 1)
    E = pos_records(str(tmp_path / "m-sample1E.pos"))
    en = pos_records(str(tmp_path / "m-sample1energy.pos"))
    assert len(E) == k.n_elems * 8 and E[0][0] == "VS" and en[0][0] == "SS"
    S = h.sample_lattice(2)
    # first sub-tetrahedron: lattice corners (0,0,0) (1,0,0) (0,1,0) (0,0,1) = indices 0, 1, 3, 6
    assert np.abs(E[0][2].reshape(4, 3) - S[0][:3, [0, 1, 3, 6]].T).max() <= 1e-9 * max(1.0, np.abs(S).max())
    e = 0.5 * (S[0][:, [0, 1, 3, 6]] ** 2).sum(axis=0)
    assert np.abs(en[0][2] - e).max() <= 1e-9


def test_norms_totals_and_errors():
    """dg_norms: totals and energies against closed forms and invariants, error
    norms against closed forms of the Gaussian pulse"""
    import dgpar
    announce("norms and totals")
    # totals of a projected constant: c * area; energy: c^2 * area
    vids, xyz, bc = mesh_tri(8, 8, lx=2.0, ly=3.0, jitter=0.1, seed=3)
    g = dgpar.DGAnalysis(dgpar.Mesh(2, vids, xyz, bc), dgpar.EulerLaw(dgpar.FLUX_LF, farfield=(1.0, 1.0, 1.0, 3.5)), 3)
    U = np.zeros(g.ndof).reshape(-1, 4, 10)
    c = np.array([1.3, -0.2, 0.7, 2.9])
    U[:, :, 0] = c / np.sqrt(2.0)
    g.set_state(U.ravel())
    n = g.norms()
    assert np.abs(n["totals"] - 6.0 * c).max() <= 1e-13 * 6.0 * np.abs(c).max()
    assert np.abs(n["energy"] - 6.0 * c * c).max() <= 1e-13 * 6.0 * (c * c).max()
    # the energy of a general state is its L2 norm squared, measured with an
    # independent collapsed Gauss rule on the oracle's basis
    m, h = vortex(p=2, n=4)
    U = h.get_state().reshape(-1, 4, 6)
    from helpers import collapsed_rule
    pts, wts = collapsed_rule(5)
    B = np.array([po.basis(2, 2, x)[0] for x in pts])  # [points][modes]
    X = m.xyz
    dj = np.abs((X[:, 1, 0] - X[:, 0, 0]) * (X[:, 2, 1] - X[:, 0, 1]) - (X[:, 2, 0] - X[:, 0, 0]) * (X[:, 1, 1] - X[:, 0, 1]))
    vals = np.einsum("efk,qk->efq", U, B)
    ref = np.einsum("efq,q,e->f", vals ** 2, wts, dj)
    assert np.abs(h.norms()["energy"] - ref).max() <= 1e-12 * ref.max()
    # conservation: periodic mesh, totals unchanged by RK4 steps
    t0 = h.norms()["totals"]
    dt = h.computeDt(0.1)
    h.advanceInTime(0.0, dt, scheme=dgpar.RK4, nsteps=10)
    assert np.abs(h.norms()["totals"] - t0).max() <= 1e-12 * np.abs(t0).max()
    # error norms of the zero state against the Gaussian pulse exp(-r^2/0.01) on
    # the unit square: int g = 0.01 pi, int g^2 = 0.005 pi (tails ~e^-25)
    vids, xyz, bc = mesh_tri(24, 24)
    s = dgpar.DGAnalysis(dgpar.Mesh(2, vids, xyz, bc), dgpar.ScalarLaw((1.0, 0.5)), 4)
    s.set_state(np.zeros(s.ndof))
    e = s.norms(0.0, dgpar.FIELD_GAUSSIAN)
    assert abs(e["l1"][0] - 0.01 * np.pi) <= 1e-5
    assert abs(e["l2"][0] - np.sqrt(0.005 * np.pi)) <= 1e-5
    assert 0.9 < e["linf"][0] <= 1.0
    s.project(dgpar.FIELD_GAUSSIAN)
    e = s.norms(0.0, dgpar.FIELD_GAUSSIAN)
    assert e["l2"][0] == s.l2_error(dgpar.FIELD_GAUSSIAN)[0]
    assert e["l1"][0] <= e["l2"][0] and e["l2"][0] <= e["linf"][0]  # unit square: |.|_1 <= |.|_2 <= |.|_inf
    with pytest.raises(dgpar.DGError, match="exact field"):
        s._check(dgpar._norms(s._h, 0.0, -1, None, (ctypes.c_double * 1)(), None, None, None, None))
