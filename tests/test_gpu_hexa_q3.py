"""GPU parity of continuous Lagrange Q3 on hexahedra (64 dofs per cell, 4^3
Gauss points, the reference's Cube basis hypercube.jl:56-92 and numbering
dofhandler.jl:37-156): FP64 element kernel and row gather against the oracle.

Tolerances: every matrix entry within 1e-12 * max|A_oracle|, every vector
entry within 1e-12 * max|b_oracle|; the CSC pattern bit-exact.  FP32 has no
matrix-core instantiation for this element and is refused.
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


def _mesh(nx, ny, nz, seed=0):
    m = B.hexa_mesh(nx + 1, ny + 1, nz + 1)
    B.jitter_nodes(m, 0.15, 1.0 / max(nx, ny, nz), seed=seed)
    return m


@pytest.mark.parametrize("dims", [(3, 3, 3), (4, 2, 3)])
def test_q3_hexa_stiffness_mass_vs_oracle(dims):
    m = _mesh(*dims, seed=sum(dims))
    eta = np.random.default_rng(5).uniform(0.5, 2.0, m.ncells)
    U = B.TrialFESpace(B.FunctionSpace("Lagrange", 3), m)
    dΩ = B.Measure(B.CellDomain(m), 6)
    ηc = B.MeshCellData(eta)
    plan = B.BilinearPlan(lambda u, v: B.integrate(ηc * B.dot(B.grad(u), B.grad(v)), dΩ)
                          + B.integrate(u * v, dΩ), U, B.TestFESpace(U))
    colptr, rowval, nzval, nd = O.assemble_bilinear(_om(m), 3, ["stiffness", "mass"], eta=eta,
                                                    degree=6)
    c, r = plan.csc_structure()
    assert np.array_equal(c + 1, colptr) and np.array_equal(r + 1, rowval)
    v1 = plan.assemble_host()
    v2 = plan.assemble_host()
    err = np.abs(v1 - nzval).max() / np.abs(nzval).max()
    print(f"hexa Q3 {dims}: ndofs {nd}, nnz {nzval.size}, rel err {err:.2e}", flush=True)
    assert err <= RTOL
    assert np.array_equal(v1, v2)


def test_q3_hexa_linear_form_vs_oracle():
    m = _mesh(3, 3, 3, seed=7)
    V = B.TestFESpace(B.TrialFESpace(B.FunctionSpace("Lagrange", 3), m))
    dΩ = B.Measure(B.CellDomain(m), 6)
    b = B.assemble_linear(lambda v: B.integrate(B.PhysicalFunction(lambda x: x[:, 0]) * v, dΩ), V)
    ref = O.assemble_linear(_om(m), 3, "x", degree=6)
    err = np.abs(b - ref).max() / np.abs(ref).max()
    print(f"hexa Q3 ∫x v: rel err {err:.2e}", flush=True)
    assert err <= RTOL


def test_q3_hexa_fp32_refused():
    m = _mesh(2, 2, 2)
    U = B.TrialFESpace(B.FunctionSpace("Lagrange", 3), m)
    dΩ = B.Measure(B.CellDomain(m), 6)
    with pytest.raises(BcubeHipError, match="FP32"):
        B.assemble_bilinear(lambda u, v: B.integrate(u * v, dΩ), U, B.TestFESpace(U), mode="fp32")
