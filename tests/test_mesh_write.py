"""Gmsh mesh export (dg_mesh_write_gmsh, the mesh half of DGAnalysis::write:
mImportExport::exportDGFile, mImportExport.cc:611-642), on the CPU.

Every written mesh is read back with dg_mesh_read_gmsh and must give the same
elements: the same vertex ids (so a restart file written over the first mesh
matches the second), the same coordinates to the bit (17 digits), the same faces
as a set (from the oracle's connectivity) and the same boundary tags.  Periodic
meshes are written as separate nodes per periodic image plus 510 / 610 lines,
which the reader joins again."""
import os

import numpy as np
import pytest

from oracle import pyoracle as po

HERE = os.path.dirname(os.path.abspath(__file__))
MESHES = os.path.join(HERE, "golden", "meshes")


def faces(m):
    dim = m.dim
    o = po.Oracle(dim, m.vids, m.xyz, 1, po.LAW_SCALAR if dim == 2 else po.LAW_MAXWELL3D, po.FLUX_UPWIND,
                  **({"vel": (1.0, 0.5), "farfield": (0.0,)} if dim == 2 else {}))
    return o.faces()


def same_mesh(a, b):
    assert a.dim == b.dim and a.vids.shape == b.vids.shape
    assert (a.vids == b.vids).all()
    assert (a.xyz == b.xyz).all()
    fa, fb = faces(a), faces(b)
    assert {tuple(r) for r in fa.tolist()} == {tuple(r) for r in fb.tolist()}
    bnd = fa[fa[:, 2] < 0]
    ta = a.bc[bnd[:, 0], bnd[:, 1]] if a.bc is not None else np.full(len(bnd), 1)
    assert (b.bc[bnd[:, 0], bnd[:, 1]] == ta).all()


@pytest.mark.parametrize("px,py", [(False, False), (True, False), (False, True), (True, True)])
def test_structured_meshes_round_trip(tmp_path, px, py):
    import dgpar
    m = dgpar.mesh_tri_structured(7, 5, x0=-1.0, lx=2.0, ly=1.5, periodic_x=px, periodic_y=py, jitter=0.1, seed=9,
                                  side_tags=(40000, 1, 20000, 2))
    p = tmp_path / "m.msh"
    dgpar.mesh_write_gmsh(m, p)
    r, info = dgpar.mesh_read_gmsh(p)
    assert info["n_periodic_pairs"] == 7 * py + 5 * px
    same_mesh(m, r)
    # periodic sides are interior faces again
    f = faces(r)
    assert (f[:, 2] < 0).sum() == (0 if px else 2 * 5) + (0 if py else 2 * 7)


def test_tetrahedra_round_trip(tmp_path):
    import dgpar
    m = dgpar.mesh_tet_structured(3, 2, 2, 1.0, 2.0, 1.5)
    p = tmp_path / "t.msh"
    dgpar.mesh_write_gmsh(m, p)
    r, info = dgpar.mesh_read_gmsh(p)
    assert info["dim"] == 3 and info["n_elems"] == 72
    m.bc = np.full(m.vids.shape, 1, np.int32)  # untagged faces without neighbour read back as far field
    same_mesh(m, r)


def test_reference_mesh_round_trip(tmp_path):
    import dgpar
    m, _ = dgpar.mesh_read_gmsh(os.path.join(MESHES, "advection256tr.msh"), split_quads=True)
    p = tmp_path / "a.msh"
    dgpar.mesh_write_gmsh(m, p)
    r, info = dgpar.mesh_read_gmsh(p)
    assert info["n_quads_split"] == 0
    same_mesh(m, r)
    m, _ = dgpar.mesh_read_gmsh(os.path.join(MESHES, "airfoil2d.msh"))
    dgpar.mesh_write_gmsh(m, p)
    r, _ = dgpar.mesh_read_gmsh(p)
    same_mesh(m, r)


def test_gmsh_vertex_ids_are_node_numbers_minus_one(tmp_path):
    import dgpar
    # node numbers 5, 9, 12, 20: two triangles sharing the edge (9, 12)
    p = tmp_path / "n.msh"
    p.write_text("$NOD\n4\n5 0 0 0\n9 1 0 0\n12 0 1 0\n20 1 1 0\n$ENDNOD\n$ELM\n2\n"
                 "1 2 100 1 3 5 9 12\n2 2 100 1 3 9 20 12\n$ENDELM\n")
    m, _ = dgpar.mesh_read_gmsh(p)
    assert m.vids.tolist() == [[4, 8, 11], [8, 19, 11]]


def test_write_rejects_what_it_cannot_express(tmp_path):
    import dgpar
    m = dgpar.mesh_tri_structured(3, 3)
    bad = dgpar.Mesh(2, m.vids - 1, m.xyz, m.bc)
    with pytest.raises(dgpar.DGError, match="vertex ids"):
        dgpar.mesh_write_gmsh(bad, tmp_path / "x.msh")
    with pytest.raises(dgpar.DGError, match="cannot write"):
        dgpar.mesh_write_gmsh(m, tmp_path / "no" / "such" / "dir.msh")
    # a face pair joined by a vertex id across a diagonal translation
    v, x = m.vids.copy(), m.xyz.copy()
    x[:, :, 1] += 0.25 * x[:, :, 0]  # shear: the x-periodic images are no longer an x translation
    with pytest.raises(dgpar.DGError, match="translation"):
        dgpar.mesh_write_gmsh(dgpar.Mesh(2, v, x, m.bc), tmp_path / "s.msh")


def test_lattice_point_counts():
    import dgpar
    assert [dgpar.lattice_points(2, n) for n in (1, 2, 3, 4, 8)] == [3, 6, 10, 15, 45]
    assert [dgpar.lattice_points(3, n) for n in (1, 2, 3)] == [4, 10, 20]
    assert dgpar.lattice_points(2, 0) == 0 and dgpar.lattice_points(3, 9) == 0 and dgpar.lattice_points(1, 2) == 0
