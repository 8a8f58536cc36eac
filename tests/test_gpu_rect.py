"""GPU parity of assembly with distinct trial and test spaces (U != V on one
mesh, assemble_bilinear(a, U, V) of assembler.jl:172-186 returning an
ndofs(V) x ndofs(U) matrix) and of linear forms whose integrand holds an FE
function (materialised at the quadrature points on the device).

Tolerances: FP64, every entry within 1e-12 * max|oracle|; the CSC pattern
bit-exact.  The reference's known answers (test_assembler.jl:40-68, 241-269)
within 1e-14.
"""
import numpy as np
import pytest

import bcubehip as B
from bcubehip._lib import BcubeHipError
from oracle import oracle as O

pytestmark = pytest.mark.gpu
RTOL = 1e-12


def _om(m):
    return O.OMesh(m.dim, m.coords, m.celltypes, m.c2n_ptr, m.c2n)


def _mesh(kind, n, seed=0):
    if kind == "Hexa8":
        m = B.hexa_mesh(n + 1, n + 1, n + 1)
    else:
        m = B.rectangle_mesh(n + 1, n + 1)
        if kind == "Tri3":
            q = m.c2n.reshape(-1, 4)
            tris = np.concatenate([q[:, [0, 1, 2]], q[:, [0, 2, 3]]], axis=1).reshape(-1, 3)
            m = B.Mesh(m.nodes(), ["Tri3"] * len(tris), tris)
    B.jitter_nodes(m, 0.15, 1.0 / n, seed=seed)
    return m


CASES = [("Quad4", 5, 1, 2, 4), ("Quad4", 5, 2, 1, 4), ("Quad4", 4, 1, 2, 3),
         ("Tri3", 5, 1, 2, 4), ("Tri3", 5, 2, 1, 4), ("Tri3", 4, 2, 1, 3),
         ("Hexa8", 3, 1, 2, 4), ("Hexa8", 3, 2, 1, 4)]


@pytest.mark.parametrize("kind,n,pu,pv,deg", CASES)
def test_rect_stiffness_mass_vs_oracle(kind, n, pu, pv, deg):
    m = _mesh(kind, n, seed=n + pu)
    eta = np.random.default_rng(2).uniform(0.5, 2.0, m.ncells)
    U = B.TrialFESpace(B.FunctionSpace("Lagrange", pu), m)
    V = B.TestFESpace(B.TrialFESpace(B.FunctionSpace("Lagrange", pv), m))
    dΩ = B.Measure(B.CellDomain(m), deg)
    ηc = B.MeshCellData(eta)
    A = B.assemble_bilinear(lambda u, v: B.integrate(ηc * B.dot(B.grad(u), B.grad(v)), dΩ)
                            + B.integrate(u * v, dΩ), U, V)
    colptr, rowval, nzval, shape = O.assemble_bilinear_rect(_om(m), pu, pv, ["stiffness", "mass"],
                                                            eta=eta, degree=deg)
    assert A.shape == shape == (V.get_ndofs(), U.get_ndofs())
    assert np.array_equal(A.indptr + 1, colptr) and np.array_equal(A.indices + 1, rowval)
    err = np.abs(A.data - nzval).max() / np.abs(nzval).max()
    print(f"{kind} n={n} U=P{pu} V=P{pv} degree {deg}: {shape}, rel err {err:.2e}", flush=True)
    assert err <= RTOL


def test_rect_repeatable_and_fp32_refused():
    m = _mesh("Quad4", 4)
    U = B.TrialFESpace(B.FunctionSpace("Lagrange", 1), m)
    V = B.TestFESpace(B.TrialFESpace(B.FunctionSpace("Lagrange", 2), m))
    dΩ = B.Measure(B.CellDomain(m), 4)
    plan = B.BilinearPlan(lambda u, v: B.integrate(B.dot(B.grad(u), B.grad(v)), dΩ), U, V)
    assert np.array_equal(plan.assemble_host(), plan.assemble_host())
    with pytest.raises(BcubeHipError, match="FP64"):
        plan.assemble_host(mode="fp32")


def test_p2_p1_known_answer():
    """test_assembler.jl:241-269 through the product."""
    m = B.one_cell_mesh(":triangle", xmin=0, xmax=1, ymin=0, ymax=1)
    U = B.TrialFESpace(B.FunctionSpace("Lagrange", 2), m)
    V = B.TestFESpace(B.TrialFESpace(B.FunctionSpace("Lagrange", 1), m))
    u = B.FEFunction(U, np.arange(1.0, 7.0))
    dΩ = B.Measure(B.CellDomain(m), 3)
    b = B.assemble_linear(lambda v: B.integrate(u * v, dΩ), V)
    print("P2-P1:", b, flush=True)
    assert np.allclose(b, [97 / 120, 4 / 5, 107 / 120], rtol=0, atol=1e-14)


@pytest.mark.parametrize("kind,pu,pv,deg", [("Quad4", 2, 1, 4), ("Quad4", 1, 2, 4),
                                            ("Tri3", 2, 1, 4), ("Tri3", 2, 2, 4),
                                            ("Hexa8", 2, 1, 4)])
def test_fe_function_linear_form_vs_oracle(kind, pu, pv, deg):
    m = _mesh(kind, 3, seed=5)
    U = B.TrialFESpace(B.FunctionSpace("Lagrange", pu), m)
    V = B.TestFESpace(B.TrialFESpace(B.FunctionSpace("Lagrange", pv), m))
    vals = np.random.default_rng(6).normal(size=U.get_ndofs())
    u = B.FEFunction(U, vals)
    dΩ = B.Measure(B.CellDomain(m), deg)
    b = B.assemble_linear(lambda v: B.integrate(2.5 * u * v, dΩ), V)
    ref = 2.5 * O.assemble_linear_fe(_om(m), pu, vals, pv, "uv", degree=deg)
    err = np.abs(b - ref).max() / np.abs(ref).max()
    print(f"{kind} ∫u v, u P{pu}, v P{pv}: rel err {err:.2e}", flush=True)
    assert err <= RTOL


def test_q1_transport_known_answer():
    """test_assembler.jl:40-68 through the product: one quad [-2,2] x [-3,3],
    discontinuous Q1, u = α, ∫((c u)⋅∇v)dΩ."""
    alpha, c, dx, dy = 10.0, np.array([0.5, 1.0]), 4.0, 6.0
    m = B.one_cell_mesh(":quad")
    m.set_nodes(m.nodes() * [2.0, 3.0])
    U = B.TrialFESpace(B.FunctionSpace("Lagrange", 1), m, isContinuous=False)
    V = B.TestFESpace(U)
    dΩ = B.Measure(B.CellDomain(m), 3)
    u = B.FEFunction(U, alpha * np.ones(V.get_ndofs()))
    b = B.assemble_linear(lambda v: B.integrate(B.dot(c * u, B.grad(v)), dΩ), V)
    ref = [-alpha / 2 * (c[0] * dy + c[1] * dx), alpha / 2 * (c[0] * dy - c[1] * dx),
           alpha / 2 * (-c[0] * dy + c[1] * dx), alpha / 2 * (c[0] * dy + c[1] * dx)]
    print("Q1 transport:", b, flush=True)
    assert np.allclose(b, ref, rtol=1e-14, atol=0)


@pytest.mark.parametrize("kind,p,cont", [("Quad4", 1, False), ("Quad4", 2, True),
                                         ("Tri3", 2, True), ("Hexa8", 1, True)])
def test_transport_linear_form_vs_oracle(kind, p, cont):
    m = _mesh(kind, 3, seed=8)
    U = B.TrialFESpace(B.FunctionSpace("Lagrange", p), m, isContinuous=cont)
    V = B.TestFESpace(U)
    c = np.array([1.0, 0.5, 0.25])[:m.dim]
    vals = np.random.default_rng(4).normal(size=U.get_ndofs())
    u = B.FEFunction(U, vals)
    dΩ = B.Measure(B.CellDomain(m), 2 * p + 1)
    b = B.assemble_linear(lambda v: B.integrate(B.dot(c * u, B.grad(v)), dΩ), V)
    ref = O.assemble_linear_fe(_om(m), p, vals, p, "cu_grad", c=c, degree=2 * p + 1,
                               continuous=cont)
    err = np.abs(b - ref).max() / np.abs(ref).max()
    print(f"{kind} P{p} continuous={cont} ∫(c u)⋅∇v: rel err {err:.2e}", flush=True)
    assert err <= RTOL
    with pytest.raises(BcubeHipError, match="FP64"):
        B.assemble_linear(lambda v: B.integrate(B.dot(c * u, B.grad(v)), dΩ), V, mode="fp32")


@pytest.mark.parametrize("kind,pu,pv,deg", [("Quad4", 1, 1, 2), ("Quad4", 2, 2, 4), ("Quad4", 1, 2, 4),
                                            ("Tri3", 2, 2, 4), ("Tri3", 2, 1, 4), ("Hexa8", 1, 1, 2),
                                            ("Hexa8", 2, 2, 4)])
def test_advection_matrix_vs_oracle(kind, pu, pv, deg):
    """∫ α (c⋅∇u) v (not symmetric; U == V included): for w in V's space,
    Aᵀ w is the oracle's ∫ w_h (α c)⋅∇λ_i over U's basis (test_assembler.jl:40-68's
    form), with α per cell; A·1 = 0 (constants have no gradient)."""
    m = _mesh(kind, 3, seed=11)
    U = B.TrialFESpace(B.FunctionSpace("Lagrange", pu), m)
    V = U if pu == pv else B.TrialFESpace(B.FunctionSpace("Lagrange", pv), m)
    c = np.array([1.0, -0.5, 0.25])[:m.dim]
    dΩ = B.Measure(B.CellDomain(m), deg)
    A = B.assemble_bilinear(lambda u, v: B.integrate(B.dot(c, B.grad(u)) * v, dΩ), U,
                            B.TestFESpace(V))
    assert A.shape == (V.get_ndofs(), U.get_ndofs())
    w = np.random.default_rng(12).normal(size=V.get_ndofs())
    ref = O.assemble_linear_fe(_om(m), pv, w, pu, "cu_grad", c=c, degree=deg)
    err = np.abs(A.T @ w - ref).max() / np.abs(ref).max()
    one = np.abs(A @ np.ones(U.get_ndofs())).max() / abs(A).max()
    print(f"{kind} U=P{pu} V=P{pv}: advection Aᵀw rel err {err:.2e}, A·1 {one:.1e}", flush=True)
    assert err <= 1e-12 and one <= 1e-13
    # a per-cell coefficient scales each cell's contribution
    alpha = np.random.default_rng(13).uniform(0.5, 2.0, m.ncells)
    ac = B.MeshCellData(alpha)
    A2 = B.assemble_bilinear(lambda u, v: B.integrate(ac * B.dot(c, B.grad(u)) * v, dΩ), U,
                             B.TestFESpace(V))
    assert np.abs(A2 @ np.ones(U.get_ndofs())).max() <= 1e-13 * abs(A2).max()
    if pu == pv:  # U == V: repeatable, and the plan refuses FP32
        plan = B.BilinearPlan(lambda u, v: B.integrate(B.dot(c, B.grad(u)) * v, dΩ), U,
                              B.TestFESpace(V))
        assert np.array_equal(plan.assemble_host(), A.data)
        with pytest.raises(BcubeHipError, match="FP64"):
            plan.assemble_host(mode="fp32")
