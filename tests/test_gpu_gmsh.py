"""The reference's gmsh meshes (test/assets/*.msh22, copied under tests/golden/)
through the product path: bilinear assembly against the oracle on the mixed
112-triangle / 45-quad mesh (FP64 within 1e-12 * max|A|, FP32 within 1e-6,
pattern bit-exact), the reference's L2 projection test on the triangle mesh
(test_projection.jl:71-101: P1, Measure degree 2, f = x + y recovered at the
vertices), and Laplace solves with Dirichlet values on the four boundary tags
of the file reproducing a linear function on the mixed mesh."""
import os

import numpy as np
import pytest
import scipy.sparse.linalg as spla

import bcubehip as B
from oracle import oracle as O

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
TRI = os.path.join(GOLDEN, "rectangle-mesh-tri-nx3-ny4.msh22")
MIXED = os.path.join(GOLDEN, "rectangle-mesh-tri-quad-nx10-ny10.msh22")
TAGS = ("South", "East", "North", "West")


def _om(m):
    return O.OMesh(m.dim, m.coords, m.celltypes, m.c2n_ptr, m.c2n)


@pytest.mark.parametrize("p,deg", [(1, 2), (1, 3), (2, 4), (3, 6)])
def test_gmsh_mixed_mesh_vs_oracle(p, deg):
    m = B.read_mesh(MIXED, warn=False)
    kappa = np.random.default_rng(p).uniform(0.5, 2.0, m.ncells)
    U = B.TrialFESpace(B.FunctionSpace("Lagrange", p), m)
    dΩ = B.Measure(B.CellDomain(m), deg)
    kc = B.MeshCellData(kappa)
    plan = B.BilinearPlan(lambda u, v: B.integrate(kc * B.dot(B.grad(u), B.grad(v)), dΩ)
                          + B.integrate(u * v, dΩ), U, B.TestFESpace(U))
    colptr, rowval, nzval, n = O.assemble_bilinear(_om(m), p, ["stiffness", "mass"], eta=kappa,
                                                   degree=deg)
    c, r = plan.csc_structure()
    assert plan.nrows == n
    assert np.array_equal(c + 1, colptr) and np.array_equal(r + 1, rowval)
    scale = np.abs(nzval).max()
    e64 = np.abs(plan.assemble_host() - nzval).max() / scale
    e32 = np.abs(plan.assemble_host(mode="fp32") - nzval).max() / scale
    print(f"gmsh mixed P{p} deg {deg}: FP64 {e64:.2e}, FP32 {e32:.2e}")
    assert e64 <= 1e-12 and e32 <= 1e-6


def test_reference_projection_on_gmsh_triangles():
    m = B.read_mesh(TRI, warn=False)
    dΩ = B.Measure(B.CellDomain(m), 2)
    U = B.TrialFESpace(B.FunctionSpace("Lagrange", 1), m, "continuous")
    u = B.FEFunction(U)
    f = B.PhysicalFunction(lambda x: x[:, 0] + x[:, 1])
    B.projection_l2_(u, f, dΩ)
    X = B.dof_coordinates(U)
    err = np.abs(u.get_dof_values() - (X[:, 0] + X[:, 1])).max()
    print("projection of x + y on the gmsh triangles: max err", err)
    assert err <= 1e-12
    # projection_l2!(u, f, mesh): degree 2*1 + 1 = 3 (the 4-point rule)
    w = B.FEFunction(U)
    B.projection_l2_(w, f, m)
    assert np.abs(w.get_dof_values() - (X[:, 0] + X[:, 1])).max() <= 1e-12


@pytest.mark.parametrize("p", [1, 2])
def test_laplace_on_gmsh_mixed_mesh_reproduces_linear(p):
    m = B.read_mesh(MIXED, warn=False)
    lin = lambda x: 0.5 + 2.0 * x[:, 0] - x[:, 1]
    g = B.PhysicalFunction(lin)
    U = B.TrialFESpace(B.FunctionSpace("Lagrange", p), m, {t: g for t in TAGS})
    V = B.TestFESpace(U)
    dΩ = B.Measure(B.CellDomain(m), 2 * p)
    A = B.assemble_bilinear(lambda u, v: B.integrate(B.dot(B.grad(u), B.grad(v)), dΩ), U, V)
    b = np.zeros(U.get_ndofs())
    B.apply_dirichlet_to_matrix_(A, U, V, m)
    B.apply_dirichlet_to_vector_(b, U, V, m)
    x = spla.spsolve(A.tocsc(), b)
    X = B.dof_coordinates(U)
    err = np.abs(x - lin(X)).max()
    print(f"P{p} Laplace on the gmsh mixed mesh: max err {err:.2e}")
    assert err <= 1e-12


@pytest.mark.parametrize("p", [1, 2])
def test_projection_on_gmsh_mixed_mesh(p):
    m = B.read_mesh(MIXED, warn=False)
    U = B.TrialFESpace(B.FunctionSpace("Lagrange", p), m)
    u = B.FEFunction(U)
    lin = lambda x: 1.0 - 2.0 * x[:, 0] + 0.5 * x[:, 1]  # noqa: E731
    B.projection_l2_(u, B.PhysicalFunction(lin), B.Measure(B.CellDomain(m), 2 * p + 1))
    err = np.abs(u.get_dof_values() - lin(B.dof_coordinates(U))).max()
    print(f"P{p} projection on the gmsh mixed mesh: max err {err:.2e}")
    assert err <= 1e-12


def test_poisson_on_gmsh_mixed_mesh_reproduces_quadratic():
    """-Δu = -4 with u = x² + y² on the boundary: P2 triangles and Q2 quads
    hold the solution, so the Galerkin solution is it."""
    m = B.read_mesh(MIXED, warn=False)
    quad = lambda x: x[:, 0] ** 2 + x[:, 1] ** 2  # noqa: E731
    U = B.TrialFESpace(B.FunctionSpace("Lagrange", 2), m, {t: B.PhysicalFunction(quad)
                                                          for t in TAGS})
    V = B.TestFESpace(U)
    dΩ = B.Measure(B.CellDomain(m), 4)
    A = B.assemble_bilinear(lambda u, v: B.integrate(B.dot(B.grad(u), B.grad(v)), dΩ), U, V)
    b = B.assemble_linear(lambda v: B.integrate(-4.0 * v, dΩ), V)
    B.apply_dirichlet_to_matrix_(A, U, V, m)
    B.apply_dirichlet_to_vector_(b, U, V, m)
    x = spla.spsolve(A.tocsc(), b)
    err = np.abs(x - quad(B.dof_coordinates(U))).max()
    print(f"P2 Poisson on the gmsh mixed mesh: max err {err:.2e}")
    assert err <= 1e-11
