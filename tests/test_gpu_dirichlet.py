"""Dirichlet conditions through the device path (bh_dirichlet_apply_*): the
reference's own vectors (test/assembler/test_dirichlet.jl, as
tests/golden/dirichlet_cases.json), host- and device-resident values giving
the same bits, and whole-path solves that linear/bilinear elements must
reproduce exactly (u = x + y, u = 1 + x - 2y + 3z)."""
import json
import os

import numpy as np
import pytest
import scipy.sparse.linalg as spla

import bcubehip as B

pytestmark = pytest.mark.gpu
GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "dirichlet_cases.json")))
Q1 = B.FunctionSpace("Lagrange", 1)


def test_matrix_edits_match_reference_vectors():
    g = GOLD["matrix"]
    m = B.rectangle_mesh(3, 2)
    U = B.TrialFESpace(Q1, m, {"xmin": 0.0})
    V = B.TestFESpace(U)
    dΩ = B.Measure(B.CellDomain(m), 1)
    A = B.assemble_bilinear(lambda u, v: B.integrate(u * v, dΩ), U, V)
    Bm = A.copy()
    B.apply_dirichlet_to_matrix_((A, Bm), U, V, m, diag_values=tuple(g["diag_values"]))
    n = A.shape[0]
    J = np.repeat(np.arange(n), np.diff(A.indptr))
    assert np.array_equal(A.indices + 1, g["IA"]) and np.array_equal(J + 1, g["JA"])
    assert np.allclose(A.data, g["VA"], rtol=0, atol=1e-15)
    assert np.allclose(Bm.data, g["VB"], rtol=0, atol=1e-15)


@pytest.mark.parametrize("case", ["fefunction", "physical"])
def test_dirichlet_vectors_match_reference(case):
    g = GOLD[case]
    m = B.rectangle_mesh(4, 3)
    if case == "fefunction":
        u = B.FEFunction(B.TrialFESpace(Q1, m), m, B.PhysicalFunction(lambda x: x[:, 1] + 1))
        U = B.TrialFESpace(Q1, m, {"xmin": lambda t: u})
    else:
        U = B.TrialFESpace(Q1, m, {"xmin": B.PhysicalFunction(lambda x: x[:, 1])})
    x = B.assemble_dirichlet_vector(U, B.TestFESpace(U), m, 0.0)
    ref = np.zeros(g["n"])
    ref[np.asarray(g["I"]) - 1] = g["V"]
    assert np.array_equal(x, ref)


def _solve_dirichlet(U, m, deg):
    V = B.TestFESpace(U)
    dΩ = B.Measure(B.CellDomain(m), deg)
    A = B.assemble_bilinear(lambda u, v: B.integrate(B.dot(B.grad(u), B.grad(v)), dΩ), U, V)
    b = np.zeros(U.get_ndofs())
    B.apply_dirichlet_to_matrix_(A, U, V, m)
    B.apply_dirichlet_to_vector_(b, U, V, m)
    return spla.spsolve(A.tocsc(), b)


def test_laplace_q1_quads_reproduces_linear():
    m = B.rectangle_mesh(17, 17)
    B.jitter_nodes(m, 0.2, 1 / 16, seed=3)
    g = B.PhysicalFunction(lambda x: x[:, 0] + x[:, 1])
    U = B.TrialFESpace(Q1, m, {t: g for t in ("xmin", "xmax", "ymin", "ymax")})
    x = _solve_dirichlet(U, m, 2)
    X = B.dof_coordinates(U)
    err = np.abs(x - (X[:, 0] + X[:, 1])).max()
    print("Q1 linear solution max err", err)
    assert err <= 1e-12


def test_laplace_q2_hexa_reproduces_linear_and_device_path():
    m = B.hexa_mesh(6, 5, 5)
    B.jitter_nodes(m, 0.15, 0.2, seed=4)
    lin = lambda x: 1 + x[:, 0] - 2 * x[:, 1] + 3 * x[:, 2]
    g = B.PhysicalFunction(lin)
    U = B.TrialFESpace(B.FunctionSpace("Lagrange", 2), m,
                       {t: g for t in ("xmin", "xmax", "ymin", "ymax", "zmin", "zmax")})
    x = _solve_dirichlet(U, m, 4)
    X = B.dof_coordinates(U)
    err = np.abs(x - lin(X)).max()
    print("Q2 hexa linear solution max err", err)
    assert err <= 1e-11
    # device-resident values: the same bits as the host path
    import torch
    V = B.TestFESpace(U)
    dΩ = B.Measure(B.CellDomain(m), 4)
    plan = B.BilinearPlan(lambda u, v: B.integrate(B.dot(B.grad(u), B.grad(v)), dΩ), U, V)
    host = plan.assemble_host().copy()
    dp = B.DirichletPlan(U)
    dp.apply_matrix(host, 1.0, homogeneous=True)
    dev = torch.zeros(plan.nnz, dtype=torch.float64, device="cuda")
    plan.assemble_device(dev.data_ptr())
    dp.apply_matrix(dev.data_ptr(), 1.0, homogeneous=True, where=1)
    bd = torch.ones(U.get_ndofs(), dtype=torch.float64, device="cuda")
    dp.apply_vector(bd.data_ptr(), 0, where=1)
    m.backend.sync()
    assert np.array_equal(dev.cpu().numpy(), host)
    bh = np.ones(U.get_ndofs())
    dp.apply_vector(bh, 0)
    assert np.array_equal(bd.cpu().numpy(), bh)
    d, src, rows, cols, diag = dp.positions()
    assert np.all(bh[d] == 0.0) and np.sum(bh == 0.0) == d.size


def test_homogeneous_system_symmetric_with_unit_rows():
    m = B.rectangle_mesh(9, 7)
    U = B.TrialFESpace(B.FunctionSpace("Lagrange", 2), m, {"xmin": 0.0, "ymin": 0.0})
    V = B.TestFESpace(U)
    dΩ = B.Measure(B.CellDomain(m), 4)
    A = B.assemble_bilinear(lambda u, v: B.integrate(B.dot(B.grad(u), B.grad(v)), dΩ)
                            + B.integrate(u * v, dΩ), U, V)
    b = np.ones(U.get_ndofs())
    B.apply_homogeneous_dirichlet_(A, b, U, V, m)
    D = A.toarray()
    d = np.unique(B.dirichlet_dofs_values(U)[0])
    # symmetric to rounding (the element kernel forms K_ij and K_ji separately)
    assert np.abs(D - D.T).max() <= 1e-14 * np.abs(D).max()
    assert np.all(D[d][:, d] == np.eye(d.size)) and np.all(b[d] == 0.0)
    assert np.all(D[d].sum(axis=1) == 1.0)


def test_reference_heat_solver_3d_q2():
    """The reference's 'Heat solver 3d', Q2 part (test_assembler.jl:457-509):
    hexa_mesh(5,5,5) on [-0.5,0.5]^3, ∫η∇u⋅∇v = ∫q v with η = 100, q = 1500,
    T = 260 on xmin, solved as AffineFESystem.solve! does
    (affine_fe_system.jl:100-119: d = Dirichlet vector, b0 = b - A d,
    homogeneous conditions on (A, b0), x + d); the cell-centre values against
    260 + (q/η) s (1 - s/2), s = x + 1/2, within the reference's 2e-14."""
    m = B.hexa_mesh(5, 5, 5, xmin=-0.5, xmax=0.5, ymin=-0.5, ymax=0.5, zmin=-0.5, zmax=0.5)
    lam, q = 100.0, 1500.0
    U = B.TrialFESpace(B.FunctionSpace("Lagrange", 2), m, {"xmin": 260.0})
    V = B.TestFESpace(U)
    dΩ = B.Measure(B.CellDomain(m), 5)
    A = B.assemble_bilinear(lambda u, v: B.integrate(lam * B.dot(B.grad(u), B.grad(v)), dΩ), U, V)
    b = B.assemble_linear(lambda v: B.integrate(q * v, dΩ), V)
    d = B.assemble_dirichlet_vector(U, V, m)
    b0 = b - A @ d
    B.apply_homogeneous_dirichlet_(A, b0, U, V, m)
    x = spla.spsolve(A.tocsc(), b0) + d
    # var_on_centers: the Q2 value at ξ = 0 is the cell-centre dof's value
    centre = [U.get_dofs(c)[13] - 1 for c in range(m.ncells)]
    Tcn = x[centre]
    xc = np.array([m.nodes()[m.cell_nodes(c)].mean(axis=0) for c in range(m.ncells)])
    s = xc[:, 0] + 0.5
    Tca = 260.0 + (q / lam) * s * (1.0 - 0.5 * s)
    err = np.abs(Tcn - Tca).max() / np.abs(Tca).max()
    print("heat solver 3d Q2 rel err", err)
    assert err < 2.0e-14
