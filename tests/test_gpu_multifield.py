"""Vector spaces and MultiFESpace through the product path (element kernels
into scalar block buffers, bh_blocks_assemble into the global CSC) against
the oracle's block assembly (assembler.jl:327-420): pattern bit-exact, FP64
within 1e-12 * max|A|, FP32 within 1e-6; vector linear forms (f⋅v, per
component through bh_blocks_add_field) against the scalar path bit for bit;
and an L2 projection of a linear vector field reproduced exactly."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

import bcubehip as B
from oracle import oracle as O

pytestmark = pytest.mark.gpu
LAG = lambda p: B.FunctionSpace("Lagrange", p)  # noqa: E731


def _om(m):
    return O.OMesh(m.dim, m.coords, m.celltypes, m.c2n_ptr, m.c2n)


def _check(plan, ref):
    colptr, rowval, nzval, n = ref
    c, r = plan.csc_structure()
    assert plan.nrows == n
    assert np.array_equal(c + 1, colptr) and np.array_equal(r + 1, rowval)
    scale = np.abs(nzval).max()
    e64 = np.abs(plan.assemble_host() - nzval).max() / scale
    e32 = np.abs(plan.assemble_host(mode="fp32") - nzval).max() / scale
    assert e64 <= 1e-12 and e32 <= 1e-6, (e64, e32)
    return e64, e32


@pytest.mark.parametrize("kind,p,size", [("quad", 2, 2), ("mixed", 1, 2), ("mixed", 3, 2),
                                         ("hexa", 1, 3), ("hexa", 2, 3)])
def test_vector_space_vs_oracle(kind, p, size):
    if kind == "quad":
        m = B.rectangle_mesh(9, 7)
        B.jitter_nodes(m, 0.2, 1 / 8, seed=1)
    elif kind == "mixed":
        m, _ = B.mixed_tri_quad_mesh(10, block=2, p_split=0.5, seed=p)
        B.jitter_nodes(m, 0.2, 0.1, seed=p + 1)
    else:
        m = B.hexa_mesh(5, 4, 4)
        B.jitter_nodes(m, 0.15, 0.25, seed=2)
    kappa = np.random.default_rng(p).uniform(0.5, 2.0, m.ncells)
    U = B.TrialFESpace(LAG(p), m, size=size)
    dΩ = B.Measure(B.CellDomain(m), 2 * p)
    kc = B.MeshCellData(kappa)
    plan = B.MultiBilinearPlan(lambda u, v: B.integrate(kc * B.ddot(B.grad(u), B.grad(v)), dΩ)
                               + B.integrate(B.dot(u, v), dΩ), U, B.TestFESpace(U))
    ref = O.assemble_blocks(_om(m), p, [size], {(0, 0): (["stiffness", "mass"], kappa)},
                            degree=2 * p)
    e64, e32 = _check(plan, ref)
    print(f"vector {kind} P{p} x{size}: FP64 {e64:.2e}, FP32 {e32:.2e}, nnz {plan.nnz}")


@pytest.mark.parametrize("aos", [True, False])
def test_multifespace_vs_oracle(aos):
    m, _ = B.mixed_tri_quad_mesh(8, block=2, p_split=0.5, seed=4)
    B.jitter_nodes(m, 0.2, 0.125, seed=5)
    sp = [B.TrialFESpace(LAG(2), m), B.TrialFESpace(LAG(2), m, size=2), B.TrialFESpace(LAG(2), m)]
    U = B.MultiFESpace(*sp, arrayOfStruct=aos)
    V = B.MultiFESpace(*[B.TestFESpace(s) for s in sp], arrayOfStruct=aos)
    dΩ = B.Measure(B.CellDomain(m), 4)

    def a(u, v):
        return (B.integrate(u[0] * v[0], dΩ) + B.integrate(B.ddot(B.grad(u[1]), B.grad(v[1])), dΩ)
                + B.integrate(u[0] * v[2] + 2.0 * B.dot(B.grad(u[2]), B.grad(v[2])), dΩ))
    plan = B.MultiBilinearPlan(a, U, V)
    blocks = {(0, 0): (["mass"], None), (1, 1): (["stiffness"], None), (2, 0): (["mass"], None),
              (2, 2): (["stiffness"], np.full(m.ncells, 2.0))}
    e64, e32 = _check(plan, O.assemble_blocks(_om(m), 2, [1, 2, 1], blocks, aos=aos, degree=4))
    print(f"MultiFESpace aos={aos}: FP64 {e64:.2e}, FP32 {e32:.2e}")
    # the device path writes the same bits
    import torch
    dev = torch.empty(plan.nnz, dtype=torch.float64, device="cuda")
    plan.assemble_device(dev.data_ptr())
    m.backend.sync()
    assert np.array_equal(dev.cpu().numpy(), plan.assemble_host())


@pytest.mark.parametrize("aos", [True, False])
def test_fully_coupled_multifespace_vs_oracle(aos):
    """All four pairs of two scalar fields coupled (off-diagonal blocks carry
    values), against the oracle."""
    m, _ = B.mixed_tri_quad_mesh(8, block=2, p_split=0.5, seed=8)
    B.jitter_nodes(m, 0.2, 0.125, seed=9)
    sp = [B.TrialFESpace(LAG(2), m), B.TrialFESpace(LAG(2), m)]
    U = B.MultiFESpace(*sp, arrayOfStruct=aos)
    V = B.MultiFESpace(*[B.TestFESpace(s) for s in sp], arrayOfStruct=aos)
    dΩ = B.Measure(B.CellDomain(m), 4)

    def a(u, v):
        return B.integrate(u[0] * v[0] + B.dot(B.grad(u[1]), B.grad(v[1])) + u[0] * v[1]
                           + 3.0 * B.dot(B.grad(u[1]), B.grad(v[0])), dΩ)
    plan = B.MultiBilinearPlan(a, U, V)
    blocks = {(0, 0): (["mass"], None), (1, 1): (["stiffness"], None), (1, 0): (["mass"], None),
              (0, 1): (["stiffness"], np.full(m.ncells, 3.0))}
    e64, e32 = _check(plan, O.assemble_blocks(_om(m), 2, [1, 1], blocks, aos=aos, degree=4))
    print(f"fully coupled aos={aos}: FP64 {e64:.2e}, FP32 {e32:.2e}")


def test_many_fields_vs_oracle():
    """17 fields, 33 coupled pairs sharing one scalar block, against the oracle."""
    m, _ = B.mixed_tri_quad_mesh(4, block=2, p_split=0.5, seed=2)
    nf = 17
    sp = [B.TrialFESpace(LAG(1), m) for _ in range(nf)]
    U = B.MultiFESpace(*sp)
    V = B.MultiFESpace(*[B.TestFESpace(s) for s in sp])
    dΩ = B.Measure(B.CellDomain(m), 2)

    def a(u, v):
        e = B.integrate(u[0] * v[0], dΩ)
        for i in range(1, nf):
            e = e + B.integrate(u[i] * v[i] + u[i - 1] * v[i], dΩ)
        return e
    plan = B.MultiBilinearPlan(a, U, V)
    assert len(plan.blocks) == 1
    blocks = {(0, 0): (["mass"], None)}
    for i in range(1, nf):
        blocks[(i, i)] = (["mass"], None)
        blocks[(i, i - 1)] = (["mass"], None)
    _check(plan, O.assemble_blocks(_om(m), 1, [1] * nf, blocks, degree=2))


def test_vector_linear_form_matches_scalar_path():
    # linear-single-vector-fe (test_assembler.jl:95-129) on quads: f = (x, x/2)
    m = B.rectangle_mesh(6, 5)
    B.jitter_nodes(m, 0.2, 0.2, seed=3)
    U = B.TrialFESpace(LAG(2), m, size=2)
    V = B.TestFESpace(U)
    dΩ = B.Measure(B.CellDomain(m), 4)
    f = B.PhysicalFunction(lambda x: np.stack([x[:, 0], x[:, 0] / 2.0], axis=1))
    b = B.assemble_linear(lambda v: B.integrate(B.dot(f, v), dΩ), V)
    S = B.TrialFESpace(LAG(2), m)
    bs = B.assemble_linear(lambda v: B.integrate(B.PhysicalFunction(lambda x: x[:, 0]) * v, dΩ),
                           B.TestFESpace(S))
    cm = U.component_map
    assert np.array_equal(b[cm[0]], bs) and np.array_equal(b[cm[1]], bs / 2.0)
    # linear-multi-scalar-fe (test_assembler.jl:131-178): SoA numbering, two scalar fields
    W = B.MultiFESpace(S, S, arrayOfStruct=False)
    f1 = B.PhysicalFunction(lambda x: x[:, 0])
    f2 = B.PhysicalFunction(lambda x: x[:, 0] / 2.0)
    b2 = B.assemble_linear(lambda v: B.integrate(f1 * v[0] + f2 * v[1], dΩ), W)
    n = S.get_ndofs()
    assert np.array_equal(b2[:n], bs) and np.array_equal(b2[n:], bs / 2.0)


def test_vector_projection_reproduces_linear_field():
    m = B.rectangle_mesh(10, 9)  # linear forms with a PhysicalFunction need a uniform mesh
    B.jitter_nodes(m, 0.2, 1 / 9, seed=7)
    U = B.TrialFESpace(LAG(1), m, size=2)
    V = B.TestFESpace(U)
    dΩ = B.Measure(B.CellDomain(m), 2)
    lin = lambda x: np.stack([x[:, 0] + x[:, 1], 2 * x[:, 0] - x[:, 1] + 0.5], axis=1)  # noqa
    M = B.assemble_bilinear(lambda u, v: B.integrate(B.dot(u, v), dΩ), U, V)
    b = B.assemble_linear(lambda v: B.integrate(B.dot(B.PhysicalFunction(lin), v), dΩ), V)
    x = spla.spsolve(M.tocsc(), b)
    X = B.dof_coordinates(B.TrialFESpace(LAG(1), m))
    want = lin(X)
    cm = U.component_map
    err = max(np.abs(x[cm[0]] - want[:, 0]).max(), np.abs(x[cm[1]] - want[:, 1]).max())
    print("vector projection max err", err)
    assert err <= 1e-12


@pytest.mark.parametrize("form", ["mass", "stiffness"])
def test_reference_basic_bilinear_benchmark(form):
    """The reference's own assembly benchmark, bench_assemble.jl:10-28:
    rectangle_mesh(251, 11; xmax=1, ymax=0.1), Q1, TrialFESpace size 2,
    Measure degree 3, a = ∫(u ⋅ v)dΩ and m = ∫(∇(u) ⋅ ∇(v))dΩ -- against the
    oracle's block assembly."""
    m = B.rectangle_mesh(251, 11, xmax=1.0, ymax=0.1)
    U = B.TrialFESpace(LAG(1), m, size=2)
    V = B.TestFESpace(U)
    dΩ = B.Measure(B.CellDomain(m), 3)
    if form == "mass":
        a = lambda u, v: B.integrate(B.dot(u, v), dΩ)  # noqa: E731
    else:
        a = lambda u, v: B.integrate(B.dot(B.grad(u), B.grad(v)), dΩ)  # noqa: E731
    A = B.assemble_bilinear(a, U, V)
    colptr, rowval, nzval, n = O.assemble_blocks(_om(m), 1, [2], {(0, 0): ([form], None)},
                                                 degree=3)
    assert A.shape == (n, n) and n == 2 * 251 * 11
    assert np.array_equal(A.indptr + 1, colptr) and np.array_equal(A.indices + 1, rowval)
    err = np.abs(A.data - nzval).max() / np.abs(nzval).max()
    print(f"basic bilinear {form}: nnz {A.nnz}, rel err {err:.2e}")
    assert err <= 1e-12


def test_linear_multi_vector_fe_known_answer():
    """linear-multi-vector-fe (test_assembler.jl:180-212): on one_cell_mesh(:tri),
    ∫(u_xy⋅v_xy + u_p⋅v_z + ∇(u_p)⋅v_xy) over MultiFESpace(U_p, U_xy, U_z),
    A[12, 1] ≈ 1/6, A[12, 2] ≈ 1/6, A[12, 3] ≈ 1/3."""
    m = B.one_cell_mesh(":tri")
    U_xy = B.TrialFESpace(LAG(1), m, size=2)
    U_z = B.TrialFESpace(LAG(1), m)
    U_p = B.TrialFESpace(LAG(1), m)
    U = B.MultiFESpace(U_p, U_xy, U_z)
    V = B.MultiFESpace(B.TestFESpace(U_p), B.TestFESpace(U_xy), B.TestFESpace(U_z))
    dΩ = B.Measure(B.CellDomain(m), 3)

    def a1(u, v):
        (u_p, u_xy, u_z), (v_p, v_xy, v_z) = u, v
        return B.integrate(B.dot(u_xy, v_xy) + B.dot(u_p, v_z) + B.dot(B.grad(u_p), v_xy), dΩ)
    A = B.assemble_bilinear(a1, U, V).toarray()
    print(A[11, :3], flush=True)
    assert A.shape == (12, 12)
    assert np.allclose(A[11, :3], [1 / 6, 1 / 6, 1 / 3], rtol=1e-14, atol=0)
    # the ∇(u_p)⋅v_xy block: ∂φ_j/∂x_a ∫ψ_i = ∂φ_j/∂x_a |T|/3 on P1 triangles
    xy = U.get_mapping(1) - 1
    cm = U_xy.component_map
    grad = np.array([[-1.0, -1.0], [1.0, 0.0], [0.0, 1.0]]) / 2.0  # ∇φ_j on (-1,-1),(1,-1),(-1,1)
    for a in range(2):
        want = np.tile(grad[:, a], (3, 1)) * (2.0 / 3.0)
        assert np.allclose(A[np.ix_(xy[cm[a]], [0, 1, 2])], want, rtol=0, atol=1e-15)


@pytest.mark.parametrize("kind,aos", [("quad", True), ("hexa", False)])
def test_grad_scalar_dot_vector_vs_oracle(kind, aos):
    """∫(u⋅v + α ∇(p)⋅v + p q) over MultiFESpace(vector u, scalar p): the
    global pattern bit-exact with the oracle's block assembly, the diagonal
    blocks within 1e-12, no values in the (q, u) blocks, coupling rows that sum
    to zero (constants have no gradient); with a constant α the coupling block
    of component a, transposed onto w, equals the oracle's ∫ w_h α e_a⋅∇λ_j
    within 1e-12; FP32 is refused (advection terms are FP64 only)."""
    if kind == "quad":
        m = B.rectangle_mesh(7, 6)
        B.jitter_nodes(m, 0.2, 1 / 6, seed=4)
        p = 2
    else:
        m = B.hexa_mesh(4, 4, 3)
        B.jitter_nodes(m, 0.15, 0.25, seed=5)
        p = 1
    d = m.dim
    deg = 2 * p
    Uv = B.TrialFESpace(LAG(p), m, size=d)
    Up = B.TrialFESpace(LAG(p), m)
    U = B.MultiFESpace(Uv, Up, arrayOfStruct=aos)
    V = B.MultiFESpace(B.TestFESpace(Uv), B.TestFESpace(Up), arrayOfStruct=aos)
    dΩ = B.Measure(B.CellDomain(m), deg)
    ac = B.MeshCellData(np.random.default_rng(6).uniform(0.5, 2.0, m.ncells))

    def a(u, v):
        return B.integrate(B.dot(u[0], v[0]) + ac * B.dot(B.grad(u[1]), v[0]) + u[1] * v[1], dΩ)
    plan = B.MultiBilinearPlan(a, U, V)
    colptr, rowval, nzval, n = O.assemble_blocks(
        _om(m), p, [d, 1], {(0, 0): (["mass"], None), (0, 1): (["mass"], None),
                            (1, 1): (["mass"], None)}, aos=aos, degree=deg)
    c, r = plan.csc_structure()
    assert plan.nrows == n
    assert np.array_equal(c + 1, colptr) and np.array_equal(r + 1, rowval)
    A = plan.to_csc(plan.assemble_host()).toarray()
    R = np.zeros((n, n))
    R[rowval - 1, np.repeat(np.arange(n), np.diff(colptr))] = nzval
    mv, mp = U.get_mapping(0) - 1, U.get_mapping(1) - 1
    cm = Uv.component_map
    vrows = np.concatenate([mv[cm[k]] for k in range(d)])
    for blk in (np.ix_(vrows, vrows), np.ix_(mp, mp)):
        assert np.abs(A[blk] - R[blk]).max() <= 1e-12 * np.abs(R).max()
    assert np.abs(A[np.ix_(mp, vrows)]).max() == 0.0
    Cp = A[np.ix_(vrows, mp)]
    assert np.abs(Cp).max() > 0 and np.abs(Cp.sum(axis=1)).max() <= 1e-13 * np.abs(Cp).max()
    with pytest.raises(B.BcubeHipError, match="FP64"):
        plan.assemble_host(mode="fp32")
    alpha = 1.5
    plan1 = B.MultiBilinearPlan(
        lambda u, v: B.integrate(alpha * B.dot(B.grad(u[1]), v[0]), dΩ), U, V)
    A1 = plan1.to_csc(plan1.assemble_host()).toarray()
    om = _om(m)
    for k in range(d):
        w = np.random.default_rng(20 + k).normal(size=Up.get_ndofs())
        want = O.assemble_linear_fe(om, p, w, p, "cu_grad", c=alpha * np.eye(d)[k], degree=deg)
        err = np.abs(A1[np.ix_(mv[cm[k]], mp)].T @ w - want).max() / np.abs(want).max()
        print(f"{kind} P{p} aos={aos}: ∇(p)⋅v component {k} rel err {err:.2e}", flush=True)
        assert err <= 1e-12
