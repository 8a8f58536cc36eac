"""GPU parity of continuous assembly on curved cells: Quad9 and Hexa27
geometry (the quadratic maps of mapping.jl:544-570, 664-697, Jacobian, inverse
and determinant in FP64 at every point) with Lagrange Q1 and Q2 spaces.

Tolerances: FP64 every matrix entry within 1e-12 * max|A_oracle|, FP32
matrix cores within 1e-6; the CSR pattern bit-exact.  The patch test solves
the Laplace problem with linear Dirichlet data on a curved mesh of the unit
box and checks that the interior solution is that linear field (1e-10).
"""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

import bcubehip as B
from oracle import oracle as O

pytestmark = pytest.mark.gpu
RTOL = 1e-12
RTOL_FP32 = 1e-6


def _om(m):
    return O.OMesh(m.dim, m.coords, m.celltypes, m.c2n_ptr, m.c2n)


def _curved(kind, n, seed=0):
    m = (B.rectangle_mesh(n + 1, n + 1, order=2) if kind == "quad9"
         else B.hexa27_mesh(n + 1, n + 1, n + 1))
    B.sinusoidal_map(m, 0.05, seed=seed)
    return m


def _bubble(m, amp=0.04):
    """A curved mesh of the unit box: x_d += amp (1 + d) Π sin(π x_e) moves no
    boundary node."""
    X = m.nodes()
    Y = X.copy()
    b = np.prod(np.sin(np.pi * X), axis=1)
    for d in range(m.dim):
        Y[:, d] += amp * (1 + d) * b
    m.set_nodes(Y)
    return m


@pytest.mark.parametrize("kind,n,p,deg", [("quad9", 6, 1, 3), ("quad9", 6, 2, 4),
                                          ("quad9", 5, 2, 6), ("hexa27", 4, 1, 3),
                                          ("hexa27", 4, 2, 5), ("hexa27", 3, 2, 3)])
def test_curved_stiffness_mass_vs_oracle(kind, n, p, deg):
    m = _curved(kind, n, seed=1)
    eta = np.random.default_rng(2).uniform(0.5, 2.0, m.ncells)
    U = B.TrialFESpace(B.FunctionSpace("Lagrange", p), m)
    dΩ = B.Measure(B.CellDomain(m), deg)
    ηc = B.MeshCellData(eta)
    plan = B.BilinearPlan(lambda u, v: B.integrate(ηc * B.dot(B.grad(u), B.grad(v)), dΩ)
                          + B.integrate(u * v, dΩ), U, B.TestFESpace(U))
    colptr, rowval, nzval, nd = O.assemble_bilinear(_om(m), p, ["stiffness", "mass"], eta=eta,
                                                    degree=deg)
    c, r = plan.csc_structure()
    assert np.array_equal(c + 1, colptr) and np.array_equal(r + 1, rowval)
    scale = np.abs(nzval).max()
    e64 = np.abs(plan.assemble_host(mode="fp64") - nzval).max() / scale
    e32 = np.abs(plan.assemble_host(mode="fp32") - nzval).max() / scale
    print(f"{kind} n={n} Q{p} degree {deg}: fp64 {e64:.2e}, fp32 {e32:.2e}")
    assert e64 <= RTOL and 0 < e32 <= RTOL_FP32


@pytest.mark.parametrize("kind,n,p,deg", [("quad9", 6, 2, 4), ("hexa27", 4, 2, 5),
                                          ("hexa27", 4, 1, 3)])
def test_curved_linear_form_vs_oracle(kind, n, p, deg):
    """∫ x v dΩ with x sampled at the curved cells' quadrature points."""
    m = _curved(kind, n, seed=3)
    V = B.TestFESpace(B.TrialFESpace(B.FunctionSpace("Lagrange", p), m))
    dΩ = B.Measure(B.CellDomain(m), deg)
    b = B.assemble_linear(lambda v: B.integrate(B.PhysicalFunction(lambda x: x[:, 0]) * v, dΩ), V)
    ref = O.assemble_linear(_om(m), p, "x", degree=deg)
    err = np.abs(b - ref).max() / np.abs(ref).max()
    print(f"{kind} Q{p} ∫x v: rel err {err:.2e}")
    assert err <= RTOL


@pytest.mark.parametrize("kind,n", [("quad9", 6), ("hexa27", 4)])
def test_curved_patch_test_through_gpu(kind, n):
    """Isoparametric Q2 on curved cells reproduces linear fields: K u = 0 in
    the interior with u = u_lin on the boundary gives u = u_lin at every
    interior dof, with the GPU's FP64 matrix (and the FP32 matrix to 1e-5)."""
    m = _bubble((B.rectangle_mesh(n + 1, n + 1, order=2) if kind == "quad9"
                 else B.hexa27_mesh(n + 1, n + 1, n + 1)))
    U = B.TrialFESpace(B.FunctionSpace("Lagrange", 2), m)
    dΩ = B.Measure(B.CellDomain(m), 5)
    plan = B.BilinearPlan(lambda u, v: B.integrate(B.dot(B.grad(u), B.grad(v)), dΩ), U,
                          B.TestFESpace(U))
    xd = B.dof_coordinates(U)
    inner = np.all((xd > 1e-12) & (xd < 1 - 1e-12), axis=1)
    assert inner.any() and (~inner).any()
    u_lin = 1.0 + 2.0 * xd[:, 0] - 3.0 * xd[:, 1] + (0.5 * xd[:, 2] if m.dim == 3 else 0.0)
    for mode, tol in (("fp64", 1e-10), ("fp32", 1e-5)):
        K = plan.to_csc(plan.assemble_host(mode=mode)).tocsr()
        ui = spla.spsolve(K[inner][:, inner].tocsc(), -K[inner][:, ~inner] @ u_lin[~inner])
        err = np.abs(ui - u_lin[inner]).max() / np.abs(u_lin).max()
        print(f"{kind} patch test {mode}: {err:.2e}")
        assert err <= tol
