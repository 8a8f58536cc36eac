"""1D (Bar2 line meshes) through the product path: the reference's known
answers of test_assembler.jl on line meshes, and 1D element matrices against
an independent numpy assembly.

  * linear-single-scalar-fe (test_assembler.jl:3-33): ∫v and ∫x v on
    line_mesh(5), P1, closed forms;
  * linear-single-vector-fe (:104-131): ∫(f⋅v), f = (x, x/2), vector P1;
  * linear-multi-scalar-fe (:133-178): ∫(f1 v1 + f2 v2) on a MultiFESpace of
    two P1 fields (struct-of-arrays numbering);
  * Bilinear rectangular system (:233-239): ∫(u⋅v) with U = P2, V = P1 on one
    line cell, checked analytically there;
  * P0..P3 stiffness + mass on a non-uniform line mesh with a per-cell η
    against numpy's Gauss-Legendre and polynomial fits (FP64, 1e-12 of the
    largest entry), and ∫(u v) with u an FE function of another order.
"""
import numpy as np
import pytest

import bcubehip as B

pytestmark = pytest.mark.gpu
LAG = lambda p: B.FunctionSpace("Lagrange", p)  # noqa: E731


def _res_x(nx, dx):
    res = np.array([dx ** 2 * i for i in range(nx)], dtype=float)
    res[0] = dx ** 2 / 6.0
    res[-1] = dx ** 2 / 2.0 * (nx - 4.0 / 3.0)
    return res


def test_linear_single_scalar_fe():
    nx = 5
    m = B.line_mesh(nx)
    dx = 1.0 / (nx - 1)
    V = B.TestFESpace(B.TrialFESpace(LAG(1), m))
    dΩ = B.Measure(B.CellDomain(m), 3)
    b1 = B.assemble_linear(lambda v: B.integrate(v, dΩ), V)
    b2 = B.assemble_linear(lambda v: B.integrate(B.PhysicalFunction(lambda x: x[:, 0]) * v, dΩ), V)
    print("∫v:", b1, " ∫x v:", b2, flush=True)
    assert np.allclose(b1, [dx / 2, dx, dx, dx, dx / 2], rtol=0, atol=1e-16)
    assert np.allclose(b2, _res_x(nx, dx), rtol=1e-14, atol=0)


def test_linear_single_vector_and_multi_scalar_fe():
    nx = 5
    m = B.line_mesh(nx)
    dx = 1.0 / (nx - 1)
    res1 = _res_x(nx, dx)
    dΩ = B.Measure(B.CellDomain(m), 3)
    V = B.TestFESpace(B.TrialFESpace(LAG(1), m, size=2))
    f = B.PhysicalFunction(lambda x: np.stack([x[:, 0], x[:, 0] / 2.0], axis=1))
    b1 = B.assemble_linear(lambda v: B.integrate(B.dot(f, v), dΩ), V)
    # the reference's numbering: (x, x) of nodes 1-2, (y, y) of nodes 1-2, then
    # x and y alternating for the next nodes
    assert np.isclose(b1[0], res1[0]) and np.isclose(b1[1], res1[1])
    assert np.isclose(b1[2], res1[0] / 2.0) and np.isclose(b1[3], res1[1] / 2.0)
    assert np.allclose(b1[4::2], res1[2:]) and np.allclose(b1[5::2], res1[2:] / 2.0)
    S = B.TrialFESpace(LAG(1), m)
    W = B.MultiFESpace(B.TestFESpace(S), B.TestFESpace(B.TrialFESpace(LAG(1), m)),
                       arrayOfStruct=False)
    f1 = B.PhysicalFunction(lambda x: x[:, 0])
    f2 = B.PhysicalFunction(lambda x: x[:, 0] / 2.0)
    b2 = B.assemble_linear(lambda v: B.integrate(f1 * v[0] + f2 * v[1], dΩ), W)
    assert np.allclose(b2[:nx], res1) and np.allclose(b2[nx:], res1 / 2.0)


def test_bilinear_rectangular_p2_p1():
    """test_assembler.jl:233-239: U = P2, V = P1 on one_cell_mesh(:line)."""
    m = B.one_cell_mesh(":line")
    U = B.TrialFESpace(LAG(2), m)
    V = B.TestFESpace(B.TrialFESpace(LAG(1), m))
    dΩ = B.Measure(B.CellDomain(m), 2)
    A = B.assemble_bilinear(lambda u, v: B.integrate(B.dot(u, v), dΩ), U, V).toarray()
    print(A, flush=True)
    assert A.shape == (2, 3)
    assert np.allclose(A[0], [1 / 3, 2 / 3, 0.0], rtol=0, atol=1e-15)
    assert np.allclose(A[1], [0.0, 2 / 3, 1 / 3], rtol=0, atol=1e-15)


def _lagrange(p):
    xi = np.linspace(-1.0, 1.0, p + 1) if p > 0 else np.array([0.0])
    out = []
    for i in range(p + 1):
        y = np.zeros(p + 1)
        y[i] = 1.0
        out.append(np.polynomial.Polynomial.fit(xi, y, p, domain=[-1, 1], window=[-1, 1])
                   if p > 0 else np.polynomial.Polynomial([1.0]))
    return out


def _numpy_1d(x, p, eta, deg, ndofs, dofs):
    """Σ_c P_cᵀ (η_c K_c + M_c) P_c with 1D element matrices from Gauss-Legendre."""
    P = _lagrange(p)
    g, w = np.polynomial.legendre.leggauss((deg + 2) // 2)
    v = np.array([f(g) for f in P])
    d = np.array([f.deriv()(g) for f in P]) if p > 0 else np.zeros_like(v)
    A = np.zeros((ndofs, ndofs))
    for c in range(len(x) - 1):
        h = x[c + 1] - x[c]
        Ke = (d * w) @ d.T * (2.0 / h)
        Me = (v * w) @ v.T * (h / 2.0)
        A[np.ix_(dofs[c], dofs[c])] += eta[c] * Ke + Me
    return A


@pytest.mark.parametrize("p", [0, 1, 2, 3])
def test_line_stiffness_mass_vs_numpy(p):
    n = 9
    m = B.line_mesh(n)
    x = np.sort(np.concatenate([[0.0, 1.0], np.random.default_rng(p).uniform(0.05, 0.95, n - 2)]))
    m.set_nodes(x[:, None])
    eta = np.random.default_rng(10 + p).uniform(0.5, 2.0, m.ncells)
    U = B.TrialFESpace(LAG(p), m, isContinuous=p > 0)
    deg = max(2 * p, 1)
    dΩ = B.Measure(B.CellDomain(m), deg)
    ηc = B.MeshCellData(eta)
    A = B.assemble_bilinear(lambda u, v: B.integrate(ηc * B.dot(B.grad(u), B.grad(v)), dΩ)
                            + B.integrate(u * v, dΩ), U, B.TestFESpace(U)).toarray()
    dofs = [U.get_dofs(c) - 1 for c in range(m.ncells)]
    R = _numpy_1d(x, p, eta, deg, U.get_ndofs(), dofs)
    err = np.abs(A - R).max() / np.abs(R).max()
    print(f"line P{p}: ndofs {U.get_ndofs()}, rel err {err:.2e}", flush=True)
    assert err <= 1e-12


def test_line_fe_function_linear_form():
    """∫(u v) with u of order 2 and v of order 1 equals M(U=2, V=1) u."""
    m = B.line_mesh(7)
    x = np.array([0.0, 0.1, 0.25, 0.5, 0.6, 0.8, 1.0])
    m.set_nodes(x[:, None])
    U = B.TrialFESpace(LAG(2), m)
    V = B.TestFESpace(B.TrialFESpace(LAG(1), m))
    vals = np.random.default_rng(2).normal(size=U.get_ndofs())
    dΩ = B.Measure(B.CellDomain(m), 4)
    b = B.assemble_linear(lambda v: B.integrate(B.FEFunction(U, vals) * v, dΩ), V)
    M = B.assemble_bilinear(lambda u, v: B.integrate(u * v, dΩ), U, V)
    assert np.abs(b - M @ vals).max() <= 1e-14 * np.abs(b).max()


def test_bilinear_rectangular_grad_p1_p2():
    """test_assembler.jl:214-231: ∫(∇(u)⋅v) with U = P1, V = P2 on one line cell."""
    m = B.one_cell_mesh(":line")
    U = B.TrialFESpace(LAG(1), m)
    V = B.TestFESpace(B.TrialFESpace(LAG(2), m))
    dΩ = B.Measure(B.CellDomain(m), 2)
    A = B.assemble_bilinear(lambda u, v: B.integrate(B.dot(B.grad(u), v), dΩ), U, V).toarray()
    print(A, flush=True)
    assert np.allclose(A[0], [-1 / 6, 1 / 6], rtol=0, atol=1e-15)
    assert np.allclose(A[1], [-2 / 3, 2 / 3], rtol=0, atol=1e-15)
    assert np.allclose(A[2], [-1 / 6, 1 / 6], rtol=0, atol=1e-15)
