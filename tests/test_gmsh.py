"""Gmsh version-1 ingest (dg_mesh_read_gmsh; the reference's
mImportExport::importGmshFile, mImportExport.cc:77-170, and
mMesh::setPeriodicBC, mMesh.cc:457-523), on the CPU.

- a lattice written as a Gmsh file with every vertex its own node reads back as
  the structured generator's mesh: same coordinates, same faces (as a set, from
  the oracle's connectivity), same boundary tags; with the sides tagged 510 / 610
  the read mesh is the generator's periodic mesh;
- a Kuhn tetrahedral mesh written as a Gmsh file reads back with the same faces;
- the reference's own mesh files: airfoil2d.msh (2094 triangles; its 96 boundary
  edges, the airfoil and the outer square, all carry the curved-wall tag 20000) and the quadrilateral
  advection meshes split into triangles (periodic tori: no boundary faces);
- the curved wall 20000 is the 40000 wall on straight faces: the oracle's
  residual with the reference's velocity-form ghost (ALG II) equals the one with
  the momentum-form wall ghost to rounding;
- malformed files and unbuilt element types fail with a message."""
import os

import numpy as np
import pytest

from helpers import lattice_msh, write_msh
from oracle import pyoracle as po

HERE = os.path.dirname(os.path.abspath(__file__))
MESHES = os.path.join(HERE, "golden", "meshes")


def faces_set(vids, xyz, dim=2):
    o = po.Oracle(dim, vids, xyz, 1, po.LAW_SCALAR if dim == 2 else po.LAW_MAXWELL3D,
                  po.FLUX_UPWIND, **({"vel": (1.0, 0.5), "farfield": (0.0,)} if dim == 2 else {}))
    f = o.faces()
    return {tuple(r) for r in f.tolist()}, f


def boundary_bc(mesh, f):
    """tag of every face without neighbour, keyed by (element, local face)"""
    return {(int(e), int(l)): int(mesh.bc[e, l]) for e, l, r, _ in f.tolist() if r < 0}


def test_lattice_reads_back_as_the_generator_mesh(tmp_path):
    import dgpar
    p = tmp_path / "lattice.msh"
    tags = (40000, 1, 20000, 2)
    lattice_msh(p, 6, 4, side_tags=tags)
    m, info = dgpar.mesh_read_gmsh(p)
    g = dgpar.mesh_tri_structured(6, 4, periodic_x=False, periodic_y=False, side_tags=tags)
    assert info["dim"] == 2 and info["n_elems"] == 48 and info["n_nodes"] == 35 and info["n_periodic_pairs"] == 0
    assert info["n_tagged"] == 20
    assert np.abs(m.xyz - g.xyz).max() <= 1e-15
    fm, rm = faces_set(m.vids, m.xyz)
    fg, rg = faces_set(g.vids, g.xyz)
    assert fm == fg
    assert boundary_bc(m, rm) == boundary_bc(g, rg)


def test_periodic_sides_join(tmp_path):
    import dgpar
    p = tmp_path / "torus.msh"
    lattice_msh(p, 5, 3, lx=2.0, ly=1.0, side_tags=(610, 510, 610, 510))
    m, info = dgpar.mesh_read_gmsh(p)
    assert info["n_periodic_pairs"] == 5 + 3
    g = dgpar.mesh_tri_structured(5, 3, lx=2.0, ly=1.0, periodic_x=True, periodic_y=True)
    fm, rm = faces_set(m.vids, m.xyz)
    fg, _ = faces_set(g.vids, g.xyz)
    assert fm == fg
    assert (rm[:, 2] >= 0).all()  # no boundary faces left
    # x-periodic only: bottom and top stay walls
    lattice_msh(p, 5, 3, side_tags=(40000, 510, 40000, 510))
    m, info = dgpar.mesh_read_gmsh(p)
    g = dgpar.mesh_tri_structured(5, 3, periodic_x=True, periodic_y=False, side_tags=(40000, 0, 40000, 0))
    fm, rm = faces_set(m.vids, m.xyz)
    fg, rg = faces_set(g.vids, g.xyz)
    assert fm == fg and boundary_bc(m, rm) == boundary_bc(g, rg)


def test_kuhn_tetrahedra_read_back(tmp_path):
    import dgpar
    g = dgpar.mesh_tet_structured(2, 2, 2)
    # one Gmsh node per distinct coordinate, tetrahedra in the generator's order
    keys, nodes, recs = {}, [], []
    for e in range(g.vids.shape[0]):
        ids = []
        for v in range(4):
            k = tuple(np.round(g.xyz[e, v], 12))
            if k not in keys:
                keys[k] = len(nodes) + 1
                nodes.append(k)
            ids.append(keys[k])
        recs.append((4, 100, ids))
    p = tmp_path / "kuhn.msh"
    write_msh(p, nodes, recs)
    m, info = dgpar.mesh_read_gmsh(p)
    assert info["dim"] == 3 and info["n_elems"] == 48 and info["n_nodes"] == 27
    assert np.abs(m.xyz - g.xyz).max() <= 1e-12
    fm, _ = faces_set(m.vids, m.xyz, dim=3)
    fg, _ = faces_set(g.vids, g.xyz, dim=3)
    assert fm == fg


def test_reference_airfoil_mesh():
    import dgpar
    m, info = dgpar.mesh_read_gmsh(os.path.join(MESHES, "airfoil2d.msh"))
    assert info["dim"] == 2 and info["n_elems"] == 2094 and info["n_tagged"] == 96
    X = m.xyz
    area = 0.5 * ((X[:, 1, 0] - X[:, 0, 0]) * (X[:, 2, 1] - X[:, 0, 1]) - (X[:, 2, 0] - X[:, 0, 0]) * (X[:, 1, 1] - X[:, 0, 1]))
    assert (np.abs(area) > 0).all()
    _, f = faces_set(m.vids, m.xyz)
    bnd = f[f[:, 2] < 0]
    tags = m.bc[bnd[:, 0], bnd[:, 1]]
    assert bnd.shape[0] == 96 and (tags == 20000).all()
    assert 3 * 2094 == 2 * (f[:, 2] >= 0).sum() + bnd.shape[0]


@pytest.mark.parametrize("name,E,pairs", [("advection4.msh", 8, 4), ("advection16.msh", 32, 8),
                                          ("advection256tr.msh", 396 + 2 * 223, 32)])
def test_reference_quad_meshes_split(name, E, pairs):
    import dgpar
    with pytest.raises(dgpar.DGError, match="quadrilaterals"):
        dgpar.mesh_read_gmsh(os.path.join(MESHES, name))
    m, info = dgpar.mesh_read_gmsh(os.path.join(MESHES, name), split_quads=True)
    assert info["n_elems"] == E and info["n_periodic_pairs"] == pairs
    if name == "advection4.msh":
        # 2 x 2 cells on a torus: joined by vertex ids, its 9 nodes fall into 4
        # classes and edges repeat, so the mesh is read but cannot be set up
        # (periodicity by shared vertex ids needs 3 or more cells per period)
        assert len(np.unique(m.vids)) == 4
        return
    _, f = faces_set(m.vids, m.xyz)
    assert (f[:, 2] >= 0).all()  # 510 / 610 sides joined: a torus


def test_curved_wall_is_the_wall_on_straight_faces():
    import dgpar
    m, _ = dgpar.mesh_read_gmsh(os.path.join(MESHES, "airfoil2d.msh"))
    ff = (1.0, 0.5, 0.1, 2.7)
    for flux in (po.FLUX_LF, po.FLUX_ROE):
        bc40 = np.where(m.bc == 20000, 40000, m.bc)
        o20 = po.Oracle(2, m.vids, m.xyz, 2, po.LAW_EULER2D, flux, gamma=1.4, farfield=ff, bc=m.bc)
        o40 = po.Oracle(2, m.vids, m.xyz, 2, po.LAW_EULER2D, flux, gamma=1.4, farfield=ff, bc=bc40)
        rng = np.random.default_rng(flux)
        U = 0.01 * rng.standard_normal((m.vids.shape[0], 4, 6))
        U[:, :, 0] += np.array(ff) / np.sqrt(2.0)
        U = U.ravel()
        r20, _ = o20.residual(U)
        r40, _ = o40.residual(U)
        assert np.abs(r20 - r40).max() <= 1e-13 * np.abs(r40).max()
        # and the wall faces do act: a far field there gives another residual
        o1 = po.Oracle(2, m.vids, m.xyz, 2, po.LAW_EULER2D, flux, gamma=1.4, farfield=ff,
                       bc=np.where(m.bc == 20000, 1, m.bc))
        assert np.abs(o1.residual(U)[0] - r40).max() > 1e-3 * np.abs(r40).max()


def test_malformed_files(tmp_path):
    import dgpar
    with pytest.raises(dgpar.DGError, match="cannot open"):
        dgpar.mesh_read_gmsh(tmp_path / "missing.msh")
    p = tmp_path / "bad.msh"
    p.write_text("$NOD\n3\n1 0 0 0\n2 1 0 0\n3 0 1 0\n$ENDNOD\n$ELM\n1\n1 8 100 1 3 1 2 3\n$ENDELM\n")
    with pytest.raises(dgpar.DGError, match="element type 8"):
        dgpar.mesh_read_gmsh(p)
    p.write_text("$NOD\n3\n1 0 0 0\n2 1 0 0\n3 0 1 0\n$ENDNOD\n$ELM\n1\n1 2 100 1 3 1 2 9\n$ENDELM\n")
    with pytest.raises(dgpar.DGError, match="missing node"):
        dgpar.mesh_read_gmsh(p)
    p.write_text("$NOD\n3\n1 0 0 0\n2 1 0 0\n3 0 1 0\n$ENDNOD\n$ELM\n2\n1 2 100 1 3 1 2 3\n2 1 510 1 2 1 3\n$ENDELM\n")
    with pytest.raises(dgpar.DGError, match="no partner"):
        dgpar.mesh_read_gmsh(p)
    p.write_text("$NOD\n3\n1 0 0 0\n")
    with pytest.raises(dgpar.DGError):
        dgpar.mesh_read_gmsh(p)
