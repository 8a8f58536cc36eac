"""Mixed triangle/quad meshes on the host: dof numbering and the CSR pattern
of the product equal the oracle's bit for bit (no GPU needed)."""
import numpy as np
import pytest

import bcubehip as B
from oracle import oracle as O

HOST = B.BcubeBackendHostOnly()


def _mixed(n=10, seed=0):
    m, blocks = B.mixed_tri_quad_mesh(n, block=2, p_split=0.5, seed=seed, backend=HOST)
    return m, blocks


def _om(m):
    return O.OMesh(m.dim, m.coords, m.celltypes, m.c2n_ptr, m.c2n)


def test_mixed_mesh_layout():
    m, blocks = B.mixed_tri_quad_mesh(8, block=4, p_split=0.5, seed=1, backend=HOST)
    ntri = int(np.sum(m.celltypes == 2))
    nquad = int(np.sum(m.celltypes == 0))
    assert ntri % 2 == 0 and nquad + ntri // 2 == 64
    assert blocks.size == m.ncells and blocks.max() < 4
    # every cell is counter-clockwise with positive area
    X = m.nodes()
    total = 0.0
    for c in range(m.ncells):
        P = X[m.cell_nodes(c)]
        area = 0.5 * np.sum(P[:, 0] * np.roll(P[:, 1], -1) - np.roll(P[:, 0], -1) * P[:, 1])
        assert area > 0
        total += area
    assert np.isclose(total, 1.0)


@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_mixed_numbering_matches_oracle(p):
    m, _ = _mixed(seed=p)
    dh = B.DofHandler(m, B.FunctionSpace("Lagrange", p))
    ig, off, n = O.dofhandler(_om(m), p)
    assert dh.ndofs_tot == n
    assert np.array_equal(dh.iglob, ig) and np.array_equal(dh.offset, off)


@pytest.mark.parametrize("p,deg", [(1, 4), (2, 4), (3, 6), (4, 8)])
def test_mixed_pattern_bit_exact(p, deg):
    m, _ = _mixed(seed=10 + p)
    U = B.TrialFESpace(B.FunctionSpace("Lagrange", p), m)
    dΩ = B.Measure(B.CellDomain(m), deg)
    plan = B.BilinearPlan(lambda u, v: B.integrate(B.dot(B.grad(u), B.grad(v)), dΩ)
                          + B.integrate(u * v, dΩ), U, B.TestFESpace(U))
    colptr, rowval = plan.csc_structure()
    ocp, orv, _, n = O.assemble_bilinear(_om(m), p, ["mass"], degree=deg)
    assert plan.nrows == n
    assert np.array_equal(colptr + 1, ocp)
    assert np.array_equal(rowval + 1, orv)


def test_sub_mesh_keeps_nodes_and_cells():
    m, blocks = _mixed(seed=3)
    sel = np.flatnonzero(blocks % 2 == 0)
    s = B.sub_mesh(m, sel)
    assert s.ncells == sel.size and s.nnodes == m.nnodes
    for k, c in enumerate(sel[:50]):
        assert np.array_equal(s.cell_nodes(k), m.cell_nodes(c))
        assert s.celltypes[k] == m.celltypes[c]
