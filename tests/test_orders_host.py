"""Continuous spaces of per-cell order on the host: the product's numbering,
edge orders and CSR pattern equal the oracle's bit for bit (no GPU needed)."""
import numpy as np
import pytest

import bcubehip as B
from oracle import oracle as O

HOST = B.BcubeBackendHostOnly()


def _om(m):
    return O.OMesh(m.dim, m.coords, m.celltypes, m.c2n_ptr, m.c2n)


def _case(n, seed, block=2, probs=(0.4, 0.3, 0.2, 0.1), p_split=0.5):
    m, blocks = B.mixed_tri_quad_mesh(n, block=block, p_split=p_split, seed=seed, backend=HOST)
    nb = int(blocks.max()) + 1
    order = np.random.default_rng(seed + 1).choice([1, 2, 3, 4], size=nb, p=probs)[blocks]
    return m, order.astype(np.int32)


@pytest.mark.parametrize("n,seed,block", [(6, 0, 2), (10, 1, 2), (16, 2, 4), (9, 3, 3)])
def test_numbering_and_edge_orders_match_oracle(n, seed, block):
    m, order = _case(n, seed, block)
    U = B.TrialFESpace(B.FunctionSpace("Lagrange", order), m)
    dofs, eord, nd = O.order_space(_om(m), order)
    assert U.get_ndofs() == nd
    assert np.array_equal(U.cell_slots(), dofs)
    eo, ntrans = U.edge_orders()
    assert np.array_equal(eo, eord) and ntrans == int((eord.sum(axis=1) > 0).sum())
    assert ntrans > 0


@pytest.mark.parametrize("p", [1, 3])
def test_one_order_equals_the_uniform_space(p):
    m, _ = _case(8, 5)
    U = B.TrialFESpace(B.FunctionSpace("Lagrange", np.full(m.ncells, p)), m)
    V = B.TrialFESpace(B.FunctionSpace("Lagrange", p), m)
    assert U.get_ndofs() == V.get_ndofs()
    for c in range(m.ncells):
        assert np.array_equal(U.get_dofs(c), V.get_dofs(c))
    assert U.edge_orders()[1] == 0


def test_pattern_matches_oracle():
    m, order = _case(12, 7)
    U = B.TrialFESpace(B.FunctionSpace("Lagrange", order), m)
    dΩ = B.Measure(B.CellDomain(m), B.OrderDegree(4))
    plan = B.BilinearPlan(lambda u, v: B.integrate(B.dot(B.grad(u), B.grad(v)), dΩ)
                          + B.integrate(u * v, dΩ), U, B.TestFESpace(U))
    colptr, rowval = plan.csc_structure()
    ocp, orv, _, n = O.assemble_bilinear_orders(_om(m), order, ["mass"])
    assert plan.nrows == n
    assert np.array_equal(colptr + 1, ocp) and np.array_equal(rowval + 1, orv)


def test_refusals():
    m, order = _case(4, 9)
    with pytest.raises(ValueError):
        B.FunctionSpace("Lagrange", np.full(m.ncells, 5))
    with pytest.raises(NotImplementedError):
        B.TrialFESpace(B.FunctionSpace("Lagrange", order), m, "discontinuous")
    h = B.hexa_mesh(3, 3, 3, backend=HOST)
    with pytest.raises(B.BcubeHipError):
        B.TrialFESpace(B.FunctionSpace("Lagrange", np.full(h.ncells, 2)), h)
    with pytest.raises(ValueError):
        B.OrderDegree(0)
