"""gmsh 2.2 input (src/io/gmsh22_reader.jl) on the host: the reference's own
mesh assets (test/assets/*.msh22, copied as fixtures under tests/golden/)
read into the cell counts, cell types and cell order its tests state
(test_domain.jl:78-98), boundary tags checked against the node coordinates,
and the dof numbering and CSR pattern equal to the oracle's bit for bit."""
import os

import numpy as np
import pytest

import bcubehip as B
from oracle import oracle as O

HOST = B.BcubeBackendHostOnly()
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
TRI = os.path.join(GOLDEN, "rectangle-mesh-tri-nx3-ny4.msh22")
MIXED = os.path.join(GOLDEN, "rectangle-mesh-tri-quad-nx10-ny10.msh22")


def _read(path):
    return B.read_mesh(path, backend=HOST, warn=False)


def _om(m):
    return O.OMesh(m.dim, m.coords, m.celltypes, m.c2n_ptr, m.c2n)


def _areas(m):
    X = m.nodes()
    out = []
    for c in range(m.ncells):
        P = X[m.cell_nodes(c)]
        out.append(0.5 * np.sum(P[:, 0] * np.roll(P[:, 1], -1) - np.roll(P[:, 0], -1) * P[:, 1]))
    return np.array(out)


def test_triangle_mesh_layout():
    m = _read(TRI)
    assert m.dim == 2 and m.nnodes == 16 and m.ncells == 20
    assert np.all(m.celltypes == B._lib.CELLTYPE["Tri3"])
    X = m.nodes()
    lo, hi = X.min(axis=0), X.max(axis=0)
    assert np.allclose(lo, [-1.5, -1.5]) and np.allclose(hi, [-0.5, -0.5])
    a = _areas(m)
    assert np.all(a > 0) and np.isclose(a.sum(), 1.0)


def test_mixed_mesh_cells_as_the_reference_states():
    # test_domain.jl:78-98: cells 1..112 are Tri3, 113..157 Quad4
    m = _read(MIXED)
    assert m.dim == 2 and m.ncells == 157
    assert np.all(m.celltypes[:112] == B._lib.CELLTYPE["Tri3"])
    assert np.all(m.celltypes[112:] == B._lib.CELLTYPE["Quad4"])
    a = _areas(m)
    X = m.nodes()
    ext = X.max(axis=0) - X.min(axis=0)
    assert np.all(a > 0) and np.isclose(a.sum(), ext[0] * ext[1])


@pytest.mark.parametrize("path", [TRI, MIXED])
def test_boundary_tags_follow_the_coordinates(path):
    m = _read(path)
    assert list(m.bc_nodes) == ["South", "East", "North", "West"]
    X = m.nodes()
    lo, hi = X.min(axis=0), X.max(axis=0)
    side = {"South": np.isclose(X[:, 1], lo[1]), "East": np.isclose(X[:, 0], hi[0]),
            "North": np.isclose(X[:, 1], hi[1]), "West": np.isclose(X[:, 0], lo[0])}
    for name, nodes in m.bc_nodes.items():
        assert len(set(nodes.tolist())) == nodes.size
        assert set(nodes.tolist()) == set(np.flatnonzero(side[name]).tolist()), name
    # every boundary face gets a tag; their lengths add up to the perimeter
    faces = B.boundary_faces(m)
    per = 0.0
    for name, fl in faces.items():
        for c, s in fl:
            nodes = m.cell_nodes(c)
            k = len(nodes)
            p0, p1 = X[nodes[s]], X[nodes[(s + 1) % k]]
            assert side[name][nodes[s]] and side[name][nodes[(s + 1) % k]]
            per += np.linalg.norm(p1 - p0)
    assert np.isclose(per, 2 * (hi - lo).sum())


@pytest.mark.parametrize("p", [1, 2, 3])
def test_gmsh_numbering_and_pattern_match_oracle(p):
    m = _read(MIXED)
    dh = B.DofHandler(m, B.FunctionSpace("Lagrange", p))
    ig, off, n = O.dofhandler(_om(m), p)
    assert dh.ndofs_tot == n
    assert np.array_equal(dh.iglob, ig) and np.array_equal(dh.offset, off)
    U = B.TrialFESpace(B.FunctionSpace("Lagrange", p), m)
    dΩ = B.Measure(B.CellDomain(m), 2 * p)
    plan = B.BilinearPlan(lambda u, v: B.integrate(u * v, dΩ), U, B.TestFESpace(U))
    colptr, rowval = plan.csc_structure()
    ocp, orv, _, _ = O.assemble_bilinear(_om(m), p, ["mass"], degree=2 * p)
    assert np.array_equal(colptr + 1, ocp) and np.array_equal(rowval + 1, orv)


def _write(tmp_path, name, physical, nodes, elements):
    lines = ["$MeshFormat", "2.2 0 8", "$EndMeshFormat", "$PhysicalNames", str(len(physical))]
    lines += [f'{d} {t} "{n}"' for d, t, n in physical]
    lines += ["$EndPhysicalNames", "$Nodes", str(len(nodes))]
    lines += [f"{i} {x} {y} {z}" for i, (x, y, z) in nodes]
    lines += ["$EndNodes", "$Elements", str(len(elements))]
    lines += [" ".join(str(v) for v in e) for e in elements]
    lines += ["$EndElements"]
    path = tmp_path / name
    path.write_text("\n".join(lines) + "\n")
    return str(path)


def test_sparse_node_ids_are_densified_in_file_order(tmp_path):
    # node ids 10, 30, 20, 40 -> 0, 1, 2, 3 (densify, utils.jl:7-12)
    nodes = [(10, (0, 0, 0)), (30, (1, 0, 0)), (20, (1, 1, 0)), (40, (0, 1, 0))]
    els = [(1, 3, 2, 7, 1, 10, 30, 20, 40), (2, 1, 2, 3, 1, 10, 30)]
    path = _write(tmp_path, "q.msh22", [(2, 7, "Domain"), (1, 3, "South")], nodes, els)
    m = _read(path)
    assert m.dim == 2 and m.ncells == 1
    assert m.cell_nodes(0).tolist() == [0, 1, 2, 3]
    assert m.bc_nodes == {"South": m.bc_nodes["South"]} and m.bc_nodes["South"].tolist() == [0, 1]


def test_surface_in_space_keeps_three_coordinates(tmp_path):
    nodes = [(1, (0, 0, 0)), (2, (1, 0, 0)), (3, (1, 1, 1))]
    path = _write(tmp_path, "t.msh22", [(2, 1, "S")], nodes, [(1, 2, 2, 1, 1, 1, 2, 3)])
    with pytest.raises(ValueError):  # a Tri3 in 3D space: no such cell in this library
        _read(path)


def test_unassembled_cell_types_and_formats_raise(tmp_path):
    nodes = [(1, (0, 0, 0)), (2, (1, 0, 0)), (3, (0, 1, 0)), (4, (0, 0, 1))]
    path = _write(tmp_path, "tet.msh22", [(3, 1, "V")], nodes, [(1, 4, 2, 1, 1, 1, 2, 3, 4)])
    with pytest.raises(NotImplementedError):
        _read(path)
    with pytest.raises(NotImplementedError):
        B.read_mesh(str(tmp_path / "mesh.msh"), backend=HOST)
    with pytest.warns(UserWarning):
        B.read_mesh(TRI, backend=HOST)


def test_hexa27_reordering():
    # gmsh22_reader.jl:187-217: gmsh's nodes 13..16 and 17..20 change places,
    # every other type keeps its order
    g = list(range(27))
    c = B.nodes_gmsh2cgns("Hexa27", g)
    assert c[:12] == g[:12] and c[12:16] == g[16:20] and c[16:20] == g[12:16] and c[20:] == g[20:]
    assert B.nodes_gmsh2cgns("Hexa27", c) == g
    assert B.nodes_gmsh2cgns("Quad9", list(range(9))) == list(range(9))
