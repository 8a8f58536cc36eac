"""Dirichlet conditions on the host (no GPU): the dofs and values of tagged
boundaries, and the CSC positions the library computes, against the
reference's own test vectors (test/assembler/test_dirichlet.jl, copied as data
to tests/golden/dirichlet_cases.json).  The edits are replayed here in numpy at
the library's positions; the GPU tests run them through the device path."""
import json
import os

import numpy as np
import pytest
import scipy.sparse as sp

import bcubehip as B
from oracle import oracle as O

HOST = B.BcubeBackendHostOnly()
GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "dirichlet_cases.json")))
Q1 = B.FunctionSpace("Lagrange", 1)


def _om(m):
    return O.OMesh(m.dim, m.coords, m.celltypes, m.c2n_ptr, m.c2n)


def test_boundary_tags_and_faces():
    m = B.rectangle_mesh(3, 2, backend=HOST)
    f = B.boundary_faces(m)
    assert list(f) == ["xmin", "xmax", "ymin", "ymax"]
    # 2 cells: left cell side 3 is xmin, right cell side 1 is xmax, bottoms ymin, tops ymax
    assert f["xmin"] == [(0, 3)] and f["xmax"] == [(1, 1)]
    assert f["ymin"] == [(0, 0), (1, 0)] and f["ymax"] == [(0, 2), (1, 2)]
    h = B.hexa_mesh(3, 3, 3, backend=HOST)
    fh = B.boundary_faces(h)
    assert all(len(fh[t]) == 4 for t in ("xmin", "xmax", "ymin", "ymax", "zmin", "zmax"))


def test_matrix_positions_match_reference_vectors():
    g = GOLD["matrix"]
    m = B.rectangle_mesh(3, 2, backend=HOST)
    U = B.TrialFESpace(Q1, m, {"xmin": 0.0})
    dofs, vals = B.dirichlet_dofs_values(U)
    assert sorted(set(dofs.tolist())) == [0, 2] and np.all(vals == 0.0)
    colptr, rowval, nzval, n = O.assemble_bilinear(_om(m), 1, ["mass"], degree=1)
    A = sp.csc_matrix((nzval, rowval - 1, colptr - 1), shape=(n, n))
    Bm = A.copy()
    plan = B.DirichletPlan(U)
    d, src, rows, cols, diag = plan.positions()
    assert sorted(d.tolist()) == [0, 2]
    for M, dv in ((A, g["diag_values"][0]), (Bm, g["diag_values"][1])):
        M.data[rows] = 0.0
        M.data[diag] = dv
    for M, V in ((A, g["VA"]), (Bm, g["VB"])):
        Mc = M.tocsc()  # the stored structure, explicit zeros included (findnz)
        Jfull = np.repeat(np.arange(n), np.diff(Mc.indptr))
        assert np.array_equal(Mc.indices + 1, g["IA"]) and np.array_equal(Jfull + 1, g["JA"])
        assert np.array_equal(Mc.data, np.asarray(V))


def test_homogeneous_positions_cover_rows_and_columns():
    m = B.rectangle_mesh(4, 3, backend=HOST)
    U = B.TrialFESpace(Q1, m, {"xmin": 0.0, "ymax": 1.0})
    plan = B.DirichletPlan(U)
    d, src, rows, cols, diag = plan.positions()
    colptr, rowval, nzval, n = O.assemble_bilinear(_om(m), 1, ["stiffness"], degree=2)
    A = sp.csc_matrix((nzval, rowval - 1, colptr - 1), shape=(n, n))
    A.data[rows] = 0.0
    A.data[cols] = 0.0
    A.data[diag] = 1.0
    D = A.toarray()
    for k in d:
        e = np.zeros(n)
        e[k] = 1.0
        assert np.array_equal(D[k], e) and np.array_equal(D[:, k], e)
    assert np.allclose(D, D.T, atol=0)


@pytest.mark.parametrize("case", ["fefunction", "physical"])
def test_dirichlet_vector_values_match_reference(case):
    g = GOLD[case]
    m = B.rectangle_mesh(4, 3, backend=HOST)
    if case == "fefunction":
        u = B.FEFunction(B.TrialFESpace(Q1, m), m, B.PhysicalFunction(lambda x: x[:, 1] + 1))
        U = B.TrialFESpace(Q1, m, {"xmin": lambda t: u})
    else:
        U = B.TrialFESpace(Q1, m, {"xmin": B.PhysicalFunction(lambda x: x[:, 1])})
    dofs, vals = B.dirichlet_dofs_values(U, 0.0)
    x = np.zeros(g["n"])
    x[dofs] = vals
    ref = np.zeros(g["n"])
    ref[np.asarray(g["I"]) - 1] = g["V"]
    assert np.array_equal(x, ref)


def test_dof_coordinates_q2_hexa():
    m = B.hexa_mesh(3, 4, 3, backend=HOST)
    U = B.TrialFESpace(B.FunctionSpace("Lagrange", 2), m)
    X = B.dof_coordinates(U)
    # the Q2 lattice of a 2x3x2 box: every (i/4, j/6, k/4) point exactly once
    key = np.round(X * [4, 6, 4]).astype(int)
    assert len({tuple(k) for k in key}) == U.get_ndofs() == 5 * 7 * 5
