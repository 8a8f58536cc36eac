"""GPU parity of the HIP assembly path against the CPU oracle (FP64 mode).

Tolerance (FP64): every matrix entry within 1e-12 * max|A_oracle|, every
vector entry within 1e-12 * max|b_oracle|; the CSR pattern bit-exact.  At the
full 64^3 size the oracle would need ~9 GB of COO, so there the checks are
size-independent properties: K·1 = 0, Σ M = |Ω|, symmetry, determinism.
"""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

import bcubehip as B
from oracle import oracle as O

pytestmark = pytest.mark.gpu
RTOL = 1e-12


def _c2_mesh(n, seed=0):
    m = B.hexa_mesh(n + 1, n + 1, n + 1)
    B.jitter_nodes(m, 0.15, 1.0 / n, seed=seed)
    o = O.hexa_mesh(n + 1, n + 1, n + 1)
    o.coords = m.coords.copy()
    return m, o


def _c1_mesh(n=64, seed=0):
    m = B.rectangle_mesh(n + 1, n + 1)
    B.jitter_nodes(m, 0.2, 1.0 / n, seed=seed)
    o = O.rectangle_mesh(n + 1, n + 1)
    o.coords = m.coords.copy()
    return m, o


def _check_matrix(A, ref):
    colptr, rowval, nzval, n = ref
    assert A.shape == (n, n)
    assert np.array_equal(A.indptr, colptr - 1)
    assert np.array_equal(A.indices, rowval - 1)
    scale = np.abs(nzval).max()
    err = np.abs(A.data - nzval).max()
    assert err <= RTOL * scale, f"max abs err {err:.3e} vs scale {scale:.3e}"
    return err / scale


def test_c1_poisson_q1_matrix_and_rhs():
    """2D Poisson, 64x64 jittered quads, Q1, Measure(Ω,2): ∫∇u⋅∇v and ∫f v."""
    m, o = _c1_mesh()
    U = B.TrialFESpace(B.FunctionSpace("Lagrange", 1), m)
    V = B.TestFESpace(U)
    dΩ = B.Measure(B.CellDomain(m), 2)
    A = B.assemble_bilinear(lambda u, v: B.integrate(B.dot(B.grad(u), B.grad(v)), dΩ), U, V)
    print("C1 matrix rel err", _check_matrix(A, O.assemble_bilinear(o, 1, ["stiffness"], degree=2)))
    f = B.PhysicalFunction(lambda x: np.sin(np.pi * x[:, 0]) * np.sin(np.pi * x[:, 1]))
    b = B.assemble_linear(lambda v: B.integrate(f * v, dΩ), V)
    bref = O.assemble_linear(o, 1, "sinsin", degree=2)
    assert np.abs(b - bref).max() <= RTOL * np.abs(bref).max()


@pytest.mark.parametrize("n", [4, 8, 12])
def test_c2_q2_stiffness_mass_vs_oracle(n):
    """3D variable-coefficient stiffness + mass, Q2 hexa, 27-point Gauss."""
    m, o = _c2_mesh(n)
    eta = np.random.default_rng(1).uniform(0.5, 2.0, m.ncells)
    U = B.TrialFESpace(B.FunctionSpace("Lagrange", 2), m)
    V = B.TestFESpace(U)
    dΩ = B.Measure(B.CellDomain(m), 5)
    ηc = B.MeshCellData(eta)
    a = lambda u, v: B.integrate(ηc * B.dot(B.grad(u), B.grad(v)), dΩ) + B.integrate(u * v, dΩ)
    A = B.assemble_bilinear(a, U, V)
    ref = O.assemble_bilinear(o, 2, ["stiffness", "mass"], eta=eta, degree=5)
    print(f"C2 {n}^3 rel err", _check_matrix(A, ref))


@pytest.mark.parametrize("order,deg", [(2, 4), (3, 6), (4, 8)])
def test_quad_higher_order_vs_oracle(order, deg):
    m, o = _c1_mesh(n=6, seed=3)
    U = B.TrialFESpace(B.FunctionSpace("Lagrange", order), m)
    dΩ = B.Measure(B.CellDomain(m), deg)
    a = lambda u, v: B.integrate(B.dot(B.grad(u), B.grad(v)), dΩ) + B.integrate(2.5 * u * v, dΩ)
    A = B.assemble_bilinear(a, U, B.TestFESpace(U))
    K = O.assemble_bilinear(o, order, ["stiffness"], degree=deg)
    M = O.assemble_bilinear(o, order, ["mass"], degree=deg)
    ref = (K[0], K[1], K[2] + 2.5 * M[2], K[3])
    _check_matrix(A, ref)


def test_hexa_q1_vs_oracle_and_source():
    m, o = _c2_mesh(5, seed=4)
    U = B.TrialFESpace(B.FunctionSpace("Lagrange", 1), m)
    dΩ = B.Measure(B.CellDomain(m), 2)
    A = B.assemble_bilinear(lambda u, v: B.integrate(B.dot(B.grad(u), B.grad(v)), dΩ), U,
                            B.TestFESpace(U))
    _check_matrix(A, O.assemble_bilinear(o, 1, ["stiffness"], degree=2))
    b = B.assemble_linear(lambda v: B.integrate(3.0 * v, dΩ), B.TestFESpace(U))
    bref = O.assemble_linear(o, 1, "const", 3.0, degree=2)
    assert np.abs(b - bref).max() <= RTOL * np.abs(bref).max()
    assert np.isclose(b.sum(), 3.0, rtol=1e-13)


def test_heat_solver_3d_q2_through_gpu():
    """test_assembler.jl:457-509 through the GPU path: the quadratic analytic
    temperature is reproduced to 2e-14 (relative, max norm over dofs)."""
    m = B.hexa_mesh(5, 5, 5, -0.5, 0.5, -0.5, 0.5, -0.5, 0.5)
    lam, q = 100.0, 1500.0
    U = B.TrialFESpace(B.FunctionSpace("Lagrange", 2), m)
    V = B.TestFESpace(U)
    dΩ = B.Measure(B.CellDomain(m), 5)
    A = B.assemble_bilinear(lambda u, v: B.integrate(lam * B.dot(B.grad(u), B.grad(v)), dΩ), U, V)
    b = B.assemble_linear(lambda v: B.integrate(q * v, dΩ), V)
    # dof coordinates: Q2 dofs sit at the Hexa8 images of the tensor nodes
    n1 = np.array([-1.0, 0.0, 1.0])
    X = m.nodes()
    xd = np.zeros((U.get_ndofs(), 3))
    for c in range(m.ncells):
        cn = X[m.cell_nodes(c)]
        lo, hi = cn.min(axis=0), cn.max(axis=0)
        for l in range(27):
            i, j, k = l % 3, (l // 3) % 3, l // 9
            xi = np.array([n1[i], n1[j], n1[k]])
            xd[U.get_dofs(c)[l] - 1] = lo + (xi + 1) / 2 * (hi - lo)
    T = lambda x: 260.0 + (q / lam) * (x + 0.5) * (1.0 - 0.5 * (x + 0.5))
    bnd = np.isclose(xd[:, 0], -0.5)
    u = np.zeros(U.get_ndofs())
    u[bnd] = 260.0
    A = A.tocsr()
    free = ~bnd
    Aff = A[free][:, free].tocsc()
    rhs = b[free] - A[free][:, bnd] @ u[bnd]
    lu = spla.splu(Aff)
    x = lu.solve(rhs)
    for _ in range(2):  # iterative refinement: remove the direct solver's own rounding
        x = x + lu.solve(rhs - Aff @ x)
    u[free] = x
    # the reference's measure: FE solution at cell centres (var_on_centers), the
    # centre of a Q2 hexa is its local dof 13 (tensor index (1,1,1))
    centres = np.array([U.get_dofs(c)[13] - 1 for c in range(m.ncells)])
    Tc = T(xd[centres, 0])
    err = np.abs(u[centres] - Tc).max() / np.abs(Tc).max()
    err_all = np.abs(u - T(xd[:, 0])).max() / np.abs(T(xd[:, 0])).max()
    print("heat 3d rel err (centres, all dofs)", err, err_all)
    assert err < 2e-14
    assert err_all < 1e-13


def test_determinism_bitwise():
    m, _ = _c2_mesh(10, seed=5)
    eta = np.random.default_rng(2).uniform(0.5, 2.0, m.ncells)
    U = B.TrialFESpace(B.FunctionSpace("Lagrange", 2), m)
    dΩ = B.Measure(B.CellDomain(m), 5)
    ηc = B.MeshCellData(eta)
    plan = B.BilinearPlan(lambda u, v: B.integrate(ηc * B.dot(B.grad(u), B.grad(v)), dΩ)
                          + B.integrate(u * v, dΩ), U, B.TestFESpace(U))
    runs = [plan.assemble_host().copy() for _ in range(3)]
    assert all(np.array_equal(runs[0], r) for r in runs[1:])


def test_c2_full_size_invariants():
    """64^3 Q2 (135,005,697 nnz): K·1 = 0, Σ M = 1, symmetry, pattern size."""
    n = 64
    m = B.hexa_mesh(n + 1, n + 1, n + 1)
    B.jitter_nodes(m, 0.15, 1.0 / n, seed=0)
    eta = np.random.default_rng(1).uniform(0.5, 2.0, m.ncells)
    U = B.TrialFESpace(B.FunctionSpace("Lagrange", 2), m)
    V = B.TestFESpace(U)
    dΩ = B.Measure(B.CellDomain(m), 5)
    ηc = B.MeshCellData(eta)
    pk = B.BilinearPlan(lambda u, v: B.integrate(ηc * B.dot(B.grad(u), B.grad(v)), dΩ), U, V)
    assert pk.nnz == 135_005_697 and pk.nrows == 2_146_689 and pk.max_row == 125
    colptr, rowval = pk.csc_structure()
    K = pk.assemble_host()
    rows = np.add.reduceat(np.abs(K), colptr[:-1])
    rs = np.add.reduceat(K, colptr[:-1])
    assert np.abs(rs).max() <= 1e-12 * rows.max()
    pm = B.BilinearPlan(lambda u, v: B.integrate(u * v, dΩ), U, V)
    M = pm.assemble_host()
    assert abs(M.sum() - 1.0) < 1e-12
    # symmetry on a sample of entries: A[i,j] == A[j,i]
    A = pk.to_csc(K)
    rng = np.random.default_rng(7)
    idx = rng.integers(0, pk.nnz, 20000)
    cols = np.searchsorted(colptr, idx, side="right") - 1
    r = rowval[idx]
    At = np.asarray(A[cols, r]).ravel()
    assert np.abs(At - K[idx]).max() <= 1e-12 * np.abs(K).max()


# ---------------------------------------------------------------------------
# FP32 matrix-core mode: same pattern, values within 1e-6 * max|A| of the FP64
# oracle; FP64 geometry and FP64 global sum.
# ---------------------------------------------------------------------------
RTOL_FP32 = 1e-6


@pytest.mark.parametrize("n", [4, 8, 12])
def test_c2_fp32_mfma_vs_oracle(n):
    m, o = _c2_mesh(n, seed=6)
    eta = np.random.default_rng(1).uniform(0.5, 2.0, m.ncells)
    U = B.TrialFESpace(B.FunctionSpace("Lagrange", 2), m)
    dΩ = B.Measure(B.CellDomain(m), 5)
    ηc = B.MeshCellData(eta)
    a = lambda u, v: B.integrate(ηc * B.dot(B.grad(u), B.grad(v)), dΩ) + B.integrate(u * v, dΩ)
    A = B.assemble_bilinear(a, U, B.TestFESpace(U), mode="fp32")
    colptr, rowval, nzval, nd = O.assemble_bilinear(o, 2, ["stiffness", "mass"], eta=eta, degree=5)
    assert np.array_equal(A.indptr, colptr - 1) and np.array_equal(A.indices, rowval - 1)
    err = np.abs(A.data - nzval).max() / np.abs(nzval).max()
    print(f"C2 {n}^3 fp32 rel err {err:.3e}")
    assert err <= RTOL_FP32
    assert err > 0  # it is a genuinely different (FP32) computation


def test_c2_full_size_fp32_vs_fp64_gpu():
    """The headline workload at its full size, 64^3 Q2 (135,005,697 nnz), in
    FP32-MFMA: every entry within 1e-6 * max|A| of the FP64 GPU assembly (the
    oracle would need ~9 GB of COO here); the FP32 stiffness annihilates
    constants (K·1 = 0 to 1e-6 of the row sizes) and the FP32 mass sums to
    |Ω| = 1 to 1e-6."""
    import torch
    n = 64
    m = B.hexa_mesh(n + 1, n + 1, n + 1)
    B.jitter_nodes(m, 0.15, 1.0 / n, seed=0)
    eta = np.random.default_rng(1).uniform(0.5, 2.0, m.ncells)
    U = B.TrialFESpace(B.FunctionSpace("Lagrange", 2), m)
    V = B.TestFESpace(U)
    dΩ = B.Measure(B.CellDomain(m), 5)
    ηc = B.MeshCellData(eta)
    plan = B.BilinearPlan(lambda u, v: B.integrate(ηc * B.dot(B.grad(u), B.grad(v)), dΩ)
                          + B.integrate(u * v, dΩ), U, V, build="device")
    assert plan.nnz == 135_005_697
    a64 = torch.empty(plan.nnz, dtype=torch.float64, device="cuda")
    a32 = torch.empty_like(a64)
    eta_d = torch.from_numpy(eta).cuda()
    plan.assemble_device(a64.data_ptr(), [eta_d.data_ptr(), None], "fp64")
    plan.assemble_device(a32.data_ptr(), [eta_d.data_ptr(), None], "fp32")
    torch.cuda.synchronize()
    err = float((a32 - a64).abs().max() / a64.abs().max())
    print(f"C2 64^3 fp32 vs fp64 GPU: rel err {err:.3e}")
    assert 0 < err <= RTOL_FP32
    del a64, a32
    colptr, _ = plan.csc_structure()
    pk = B.BilinearPlan(lambda u, v: B.integrate(ηc * B.dot(B.grad(u), B.grad(v)), dΩ), U, V)
    K = pk.assemble_host(mode="fp32")
    rows = np.add.reduceat(np.abs(K), colptr[:-1])
    rs = np.add.reduceat(K, colptr[:-1])
    print(f"FP32 K·1: max |row sum| / max row size {np.abs(rs).max() / rows.max():.3e}")
    assert np.abs(rs).max() <= RTOL_FP32 * rows.max()
    del K, rows, rs, pk
    pm = B.BilinearPlan(lambda u, v: B.integrate(u * v, dΩ), U, V)
    M = pm.assemble_host(mode="fp32")
    print(f"FP32 Σ M - 1 = {M.sum() - 1.0:.3e}")
    assert abs(M.sum() - 1.0) <= RTOL_FP32


@pytest.mark.parametrize("order,deg", [(1, 2), (2, 4), (3, 6), (4, 8)])
def test_quad_fp32_mfma_vs_oracle(order, deg):
    m, o = _c1_mesh(n=8, seed=8)
    U = B.TrialFESpace(B.FunctionSpace("Lagrange", order), m)
    dΩ = B.Measure(B.CellDomain(m), deg)
    a = lambda u, v: B.integrate(B.dot(B.grad(u), B.grad(v)), dΩ) + B.integrate(u * v, dΩ)
    A = B.assemble_bilinear(a, U, B.TestFESpace(U), mode="fp32")
    K = O.assemble_bilinear(o, order, ["stiffness"], degree=deg)
    M = O.assemble_bilinear(o, order, ["mass"], degree=deg)
    ref = K[2] + M[2]
    assert np.abs(A.data - ref).max() <= RTOL_FP32 * np.abs(ref).max()


def test_fp32_determinism_and_solution_error():
    """Bitwise-repeatable FP32 assembly; the solution of a Q2 reaction-diffusion
    problem with the FP32 matrix, compared with the FP64-matrix solution (error
    reported; bound 1e-4, the matrix error times the conditioning)."""
    m = B.hexa_mesh(7, 7, 7)
    B.jitter_nodes(m, 0.15, 1 / 6, seed=9)
    U = B.TrialFESpace(B.FunctionSpace("Lagrange", 2), m)
    V = B.TestFESpace(U)
    dΩ = B.Measure(B.CellDomain(m), 5)
    a = lambda u, v: B.integrate(B.dot(B.grad(u), B.grad(v)), dΩ) + B.integrate(u * v, dΩ)
    p32 = B.BilinearPlan(a, U, V, mode="fp32")
    r = [p32.assemble_host().copy() for _ in range(3)]
    assert all(np.array_equal(r[0], x) for x in r[1:])
    A32 = p32.to_csc(r[0])
    A64 = B.assemble_bilinear(a, U, V)
    b = B.assemble_linear(lambda v: B.integrate(1.0 * v, dΩ), V)
    x64 = spla.spsolve(A64.tocsc(), b)
    x32 = spla.spsolve(A32.tocsc(), b)
    err = np.abs(x32 - x64).max() / np.abs(x64).max()
    print("fp32 solution rel err", err)
    # the matrix error (~5e-7 of max|A|) is amplified by the conditioning of A
    assert err <= 1e-4


def test_assemble_linear_accumulates_like_reference():
    """assemble_linear! adds into b (assembler.jl:855-865)."""
    m, o = _c1_mesh(n=8, seed=1)
    V = B.TestFESpace(B.TrialFESpace(B.FunctionSpace("Lagrange", 2), m))
    dΩ = B.Measure(B.CellDomain(m), 4)
    b = np.full(V.get_ndofs(), 2.0)
    B.assemble_linear_(b, lambda v: B.integrate(1.0 * v, dΩ), V)
    assert np.isclose(b.sum(), 2.0 * V.get_ndofs() + 1.0, rtol=1e-13)


@pytest.mark.parametrize("n", [4, 12])
def test_c2_bf16x3_split_error_reported(n):
    """BF16 hi/lo split with three products: same pattern, error ~1e-5 of max|A|
    (two 8-bit halves carry ~16 bits); the bound here is 1e-4."""
    m, o = _c2_mesh(n, seed=6)
    eta = np.random.default_rng(1).uniform(0.5, 2.0, m.ncells)
    U = B.TrialFESpace(B.FunctionSpace("Lagrange", 2), m)
    dΩ = B.Measure(B.CellDomain(m), 5)
    ηc = B.MeshCellData(eta)
    a = lambda u, v: B.integrate(ηc * B.dot(B.grad(u), B.grad(v)), dΩ) + B.integrate(u * v, dΩ)
    A = B.assemble_bilinear(a, U, B.TestFESpace(U), mode="bf16x3")
    colptr, rowval, nzval, nd = O.assemble_bilinear(o, 2, ["stiffness", "mass"], eta=eta, degree=5)
    assert np.array_equal(A.indices, rowval - 1)
    err = np.abs(A.data - nzval).max() / np.abs(nzval).max()
    print(f"C2 {n}^3 bf16x3 rel err {err:.3e}")
    assert err <= 1e-4


def test_hipgraph_replay_matches_eager():
    """Re-assembly captured as a HIP graph: every replay (with eager work and
    synchronisation between replays, after the values were clobbered) equals
    the eager assembly bitwise, for the matrix and the right-hand side."""
    import torch

    m, _ = _c2_mesh(6, seed=11)
    eta = np.random.default_rng(3).uniform(0.5, 2.0, m.ncells)
    U = B.TrialFESpace(B.FunctionSpace("Lagrange", 2), m)
    V = B.TestFESpace(U)
    dΩ = B.Measure(B.CellDomain(m), 5)
    ηc = B.MeshCellData(eta)
    plan = B.BilinearPlan(lambda u, v: B.integrate(ηc * B.dot(B.grad(u), B.grad(v)), dΩ)
                          + B.integrate(u * v, dΩ), U, V)
    dev = torch.device("cuda:0")
    backend = m.backend
    stream = torch.cuda.Stream(dev)
    backend.set_stream(stream.cuda_stream)
    eta_d = torch.from_numpy(eta).to(dev)
    out = torch.zeros(plan.nnz, dtype=torch.float64, device=dev)
    for mode in ("fp64", "fp32"):
        plan.assemble_device(out.data_ptr(), [eta_d.data_ptr(), None], mode)
        backend.sync()
        eager = out.clone()
        with backend.capture() as g:
            plan.assemble_device(out.data_ptr(), [eta_d.data_ptr(), None], mode)
        for rep in range(4):
            with torch.cuda.stream(stream):
                out.fill_(float(rep) - 7.0)  # clobber, eager work between replays
            backend.sync()
            g.launch()
            backend.sync()
            assert torch.equal(out, eager), (mode, rep)
    backend.set_stream(None)


def test_large_assembly_graph_replay_and_timing_kinds():
    """40^3 Q1 hexa (64,000 cells): the values equal the oracle's (FP64 1e-12,
    FP32 1e-6), a repeat is bitwise identical, the timed launches are one
    element kernel and one row gather, and a HIP graph of the assembly replays
    bitwise equal to the eager run."""
    import torch

    n = 40
    m, o = _c2_mesh(n, seed=13)
    U = B.TrialFESpace(B.FunctionSpace("Lagrange", 1), m)
    dΩ = B.Measure(B.CellDomain(m), 2)
    plan = B.BilinearPlan(lambda u, v: B.integrate(B.dot(B.grad(u), B.grad(v)), dΩ), U,
                          B.TestFESpace(U))
    ref = O.assemble_bilinear(o, 1, ["stiffness"], degree=2)
    backend = m.backend
    stream = torch.cuda.Stream()
    backend.set_stream(stream.cuda_stream)
    out = torch.zeros(plan.nnz, dtype=torch.float64, device="cuda")
    try:
        for mode, tol in (("fp64", RTOL), ("fp32", 1e-6)):
            plan.assemble_device(out.data_ptr(), None, mode)
            backend.sync()
            eager = out.clone()
            backend.set_timing(True)
            plan.assemble_device(out.data_ptr(), None, mode)
            ms, kinds = backend.timings_ex()
            backend.set_timing(False)
            assert list(kinds) == [0, 1] and all(t > 0 for t in ms)
            assert torch.equal(out, eager)
            A = plan.to_csc(eager.cpu().numpy())
            assert np.array_equal(A.indptr, ref[0] - 1) and np.array_equal(A.indices, ref[1] - 1)
            err = np.abs(A.data - ref[2]).max() / np.abs(ref[2]).max()
            print(mode, "rel err", err)
            assert err <= tol
            with backend.capture() as g:
                plan.assemble_device(out.data_ptr(), None, mode)
            for rep in range(3):
                with torch.cuda.stream(stream):
                    out.fill_(float(rep) + 3.0)
                backend.sync()
                g.launch()
                backend.sync()
                assert torch.equal(out, eager), (mode, rep)
    finally:
        backend.set_stream(None)


def test_c5_batched_instances_blocks_match_oracle():
    """Many-instance batch: instances of uneven size concatenated into one mesh
    per order bucket; each diagonal block of the batched matrix equals the
    oracle's matrix of that instance alone (FP64 1e-12, FP32 1e-6)."""
    rng = np.random.default_rng(3)
    sides = [int(s) for s in rng.integers(3, 9, size=6)]
    for order in (1, 2):
        inst = [B.rectangle_mesh(s + 1, s + 1) for s in sides]
        kappa = [np.exp(rng.normal(0.0, 0.5, m.ncells)) for m in inst]
        big, coff, _ = B.batch_meshes(inst)
        U = B.TrialFESpace(B.FunctionSpace("Lagrange", order), big)
        dΩ = B.Measure(B.CellDomain(big), 2 * order)
        kc = B.MeshCellData(np.concatenate(kappa))
        a = lambda u, v: B.integrate(kc * B.dot(B.grad(u), B.grad(v)), dΩ) + B.integrate(u * v, dΩ)
        for mode, tol in (("fp64", RTOL), ("fp32", RTOL_FP32)):
            A = B.assemble_bilinear(a, U, B.TestFESpace(U), mode=mode).tocsr()
            shift = 0
            for k, s in enumerate(sides):
                o = O.rectangle_mesh(s + 1, s + 1)
                K = O.csc_to_scipy(*O.assemble_bilinear(o, order, ["stiffness"], eta=kappa[k],
                                                        degree=2 * order))
                M = O.csc_to_scipy(*O.assemble_bilinear(o, order, ["mass"], degree=2 * order))
                ref = (K + M).toarray()
                n = ref.shape[0]
                blk = A[shift:shift + n, shift:shift + n].toarray()
                assert np.abs(blk - ref).max() <= tol * np.abs(ref).max(), (order, mode, k)
                # nothing couples instance k to the others
                assert A[shift:shift + n].nnz == (K + M).nnz
                shift += n
            assert shift == A.shape[0]
