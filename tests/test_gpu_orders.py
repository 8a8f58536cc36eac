"""GPU parity of assembly on continuous spaces of per-cell order (mixed
Quad4/Tri3 meshes, orders 1..4 per block of cells, edges at their cells'
lowest order): ∫η∇u⋅∇v + ∫uv with Measure degree max(2p, 4) per cell.

FP32 on a space with Q4 cells runs every (cell type, order) bucket's element
matrices in one launch, bitwise equal to one launch per bucket; the FP32
element kernels apply the constraint of the cells with an edge of lower order
in LDS, FP64 runs a constraint kernel after its element kernels.

Tolerances: FP64 every entry within 1e-12 * max|A_oracle| of the oracle's
assembly of the same space, the CSR pattern bit-exact; FP32 (FP64 geometry,
FP32 element matrices, FP64 sums) within 1e-6 * max|A_oracle|.  The solution
of (K + M) u = b with b the oracle's ∫v: within 1e-10 (FP64) and 5e-3 (FP32)
of the solution with the oracle's matrix -- the FP32 bound is the matrix
error (below 1e-6) times the system's condition number (about 1e3-1e4 at
orders up to 4 on a 40^2 mesh; measured 4.6e-4).  Repeated assemblies are bitwise
equal, and the device-coefficient call equals the host one.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import bcubehip as B
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _om(m):
    return O.OMesh(m.dim, m.coords, m.celltypes, m.c2n_ptr, m.c2n)


def _case(n, seed, block, probs=(0.4, 0.3, 0.2, 0.1), p_split=1.0 / 3.0):
    m, blocks = B.mixed_tri_quad_mesh(n, block=block, p_split=p_split, seed=seed)
    B.jitter_nodes(m, 0.2, 1.0 / n, seed=seed + 50)
    nb = int(blocks.max()) + 1
    order = np.random.default_rng(seed + 1).choice([1, 2, 3, 4], size=nb, p=probs)[blocks]
    eta = np.random.default_rng(seed + 2).uniform(0.5, 2.0, m.ncells)
    return m, order.astype(np.int32), eta


def _plan(m, order, eta, mode="fp64"):
    U = B.TrialFESpace(B.FunctionSpace("Lagrange", order), m)
    dΩ = B.Measure(B.CellDomain(m), B.OrderDegree(4))
    kc = B.MeshCellData(eta)
    form = lambda u, v: B.integrate(kc * B.dot(B.grad(u), B.grad(v)), dΩ) + B.integrate(u * v, dΩ)
    return B.BilinearPlan(form, U, B.TestFESpace(U), mode), U


def _kinds(plan, mode):
    be = plan.V.mesh.backend
    be.set_timing(True)
    try:
        vals = plan.assemble_host(mode=mode)
        _, kinds = be.timings_ex()
    finally:
        be.set_timing(False)
    return vals, list(kinds)


@pytest.mark.parametrize("n,seed,block", [(8, 0, 2), (16, 1, 4), (24, 2, 4), (30, 3, 8)])
def test_fp64_and_fp32_vs_oracle(n, seed, block):
    m, order, eta = _case(n, seed, block)
    plan, U = _plan(m, order, eta)
    cp, rv, ref, nd = O.assemble_bilinear_orders(_om(m), order, ["stiffness", "mass"], eta)
    c, r = plan.csc_structure()
    assert plan.nrows == nd
    assert np.array_equal(c + 1, cp) and np.array_equal(r + 1, rv)
    scale = np.abs(ref).max()
    v64, k64 = _kinds(plan, "fp64")
    v32, k32 = _kinds(plan, "fp32")
    e64 = np.abs(v64 - ref).max() / scale
    e32 = np.abs(v32 - ref).max() / scale
    ntrans = U.edge_orders()[1]
    print(f"n={n} orders {np.bincount(order, minlength=5)[1:]} transition cells {ntrans}: "
          f"fp64 {e64:.2e} fp32 {e32:.2e} kinds {k32}")
    assert e64 <= 1e-12
    assert 0 < e32 <= 1e-6
    # FP64: the constraint kernel (kind 4) after the element kernels; FP32: the
    # element kernel forms Cᵀ K C itself
    assert (4 in k64) == (ntrans > 0) and 4 not in k32
    assert np.array_equal(v32, plan.assemble_host(mode="fp32"))
    assert np.array_equal(v64, plan.assemble_host(mode="fp64"))
    # FP32 with Q4 cells (rows of 32 floats): every bucket's element matrices
    # in one launch, bitwise the per-bucket launches (tuning shape 12)
    q4 = bool(np.any((order == 4) & (m.celltypes == 0)))
    assert (k32.count(0) == 1) == q4, k32
    plan.set_tuning(12, mode="fp32")
    v32b, k32b = _kinds(plan, "fp32")
    assert np.array_equal(v32b, v32)
    assert k32b.count(0) == len(np.unique(np.stack([order, m.celltypes]), axis=1).T)


def test_solution_and_device_call():
    import torch

    m, order, eta = _case(40, 5, 8)
    plan, U = _plan(m, order, eta)
    om = _om(m)
    cp, rv, ref, nd = O.assemble_bilinear_orders(om, order, ["stiffness", "mass"], eta)
    b = O.assemble_linear_orders(om, order, "one")
    A_ref = O.csc_to_scipy(cp, rv, ref, nd).tocsc()
    u_ref = spla.spsolve(A_ref, b)
    c, r = plan.csc_structure()
    dev = torch.device("cuda:0")
    eta_d = torch.from_numpy(eta).to(dev)
    for mode, tol in (("fp64", 1e-10), ("fp32", 5e-3)):
        out = torch.zeros(plan.nnz, dtype=torch.float64, device=dev)
        plan.assemble_device(out.data_ptr(), [eta_d.data_ptr(), None], mode)
        m.backend.sync()
        vals = out.cpu().numpy()
        assert np.array_equal(vals, plan.assemble_host(mode=mode))
        A = sp.csc_matrix((vals, r, c), shape=(nd, nd))
        u = spla.spsolve(A, b)
        err = np.abs(u - u_ref).max() / np.abs(u_ref).max()
        print(f"{mode}: solution rel err {err:.2e}")
        assert err <= tol


def test_one_order_equals_uniform_space_on_gpu():
    """All cells of order 3: the per-cell-order space is the uniform space,
    and its matrix equals the uniform plan's bit for bit (same kernels, same
    numbering, no constrained cells)."""
    m, _, eta = _case(12, 6, 4)
    order = np.full(m.ncells, 3, dtype=np.int32)
    plan, U = _plan(m, order, eta)
    V = B.TrialFESpace(B.FunctionSpace("Lagrange", 3), m)
    dΩ = B.Measure(B.CellDomain(m), 6)
    kc = B.MeshCellData(eta)
    p2 = B.BilinearPlan(lambda u, v: B.integrate(kc * B.dot(B.grad(u), B.grad(v)), dΩ)
                        + B.integrate(u * v, dΩ), V, B.TestFESpace(V))
    for mode in ("fp64", "fp32"):
        assert np.array_equal(plan.assemble_host(mode=mode), p2.assemble_host(mode=mode))
