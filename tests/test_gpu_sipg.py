"""GPU parity of the symmetric interior-penalty Poisson forms
(benchmark/bench_assemble.jl:33-96): cell, interior-face and boundary-face
bilinear terms and the boundary-face linear term, assembled by the face
kernels + row gather, against the oracle's restatement (FP64 within
1e-12 * max|A|, CSC structure bit-exact), and the benchmark problem itself
(u = 3x + y² + 2x³ in Q3) solved exactly."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

import bcubehip as B
from oracle import oracle as O

pytestmark = pytest.mark.gpu
RTOL = 1e-12


def _rect(n, lo=-0.5, hi=0.5, jitter=0.0, seed=0, rotate=False):
    """n x n Quad4 cells on [lo, hi]², optional interior jitter and cyclic
    rotation of some cells' node lists (other local sides meet on a face)."""
    t = np.linspace(lo, hi, n + 1)
    X = np.array([[x, y] for y in t for x in t])
    if jitter:
        rng = np.random.default_rng(seed)
        inner = np.all((np.abs(X - lo) > 1e-12) & (np.abs(X - hi) > 1e-12), axis=1)
        X[inner] += rng.uniform(-jitter, jitter, (inner.sum(), 2)) * (hi - lo) / n
    cells = []
    for j in range(n):
        for i in range(n):
            a = j * (n + 1) + i
            c = [a, a + 1, a + n + 2, a + n + 1]
            if rotate and (i + j) % 2:
                k = (i + 2 * j) % 3 + 1
                c = c[k:] + c[:k]
            cells.append(c)
    return B.Mesh(X, ["Quad4"] * (n * n), np.array(cells, dtype=np.int64))


def _om(m):
    return O.OMesh(m.dim, m.coords, m.celltypes, m.c2n_ptr, m.c2n)


def _ua(x):
    return 3 * x[..., 0] + x[..., 1] ** 2 + 2 * x[..., 0] ** 3


def _f(x):
    return -2.0 - 12.0 * x[..., 0]


def _sipg(m, order, degree, pen, terms=("vol", "int", "bnd")):
    U = B.TrialFESpace(B.FunctionSpace("Lagrange", order), m, "discontinuous")
    V = B.TestFESpace(U)
    dΩ = B.Measure(B.CellDomain(m), degree)
    dΓ = B.Measure(B.InteriorFaceDomain(m), degree)
    dΓb = B.Measure(B.BoundaryFaceDomain(m), degree)
    nΓ, nΓb = B.get_face_normals(dΓ), B.get_face_normals(dΓb)

    def a(u, v):
        parts = []
        if "vol" in terms:
            parts.append(B.integrate(B.dot(B.grad(v), B.grad(u)), dΩ))
        if "int" in terms:
            parts.append(B.integrate(
                -B.dot(B.jump(v, nΓ), B.avg(B.grad(u))) - B.dot(B.avg(B.grad(v)), B.jump(u, nΓ))
                + pen * B.dot(B.jump(v, nΓ), B.jump(u, nΓ)), dΓ))
        if "bnd" in terms:
            fa = lambda u, gu, v, gv, n: -v * B.dot(gu, n) - B.dot(gv, n) * u + pen * v * u  # noqa
            parts.append(B.integrate(
                B.compose(fa, *map(B.side_n, (u, B.grad(u), v, B.grad(v), nΓb))), dΓb))
        out = parts[0]
        for p in parts[1:]:
            out = out + p
        return out

    g = B.PhysicalFunction(_ua)
    f = B.PhysicalFunction(_f)
    fl = lambda v, gv, n, g: -B.dot(gv, n) * g + pen * v * g  # noqa: E731

    def l(v):
        return (B.integrate(v * f, dΩ)
                + B.integrate(B.compose(fl, *map(B.side_n, (v, B.grad(v), nΓb, g))), dΓb))

    return U, V, a, l


def _check_matrix(A, R):
    R = R.tocsc()
    R.sum_duplicates()
    R.sort_indices()
    assert np.array_equal(A.indptr, R.indptr) and np.array_equal(A.indices, R.indices)
    err = np.abs(A.data - R.data).max() / np.abs(R.data).max()
    assert err <= RTOL, err
    return err


CASES = [(1, 4, 0.0, False), (2, 4, 0.15, True), (3, 4, 0.0, False), (3, 5, 0.2, True),
         (4, 3, 0.1, True)]


@pytest.mark.parametrize("order,n,jitter,rotate", CASES)
def test_sipg_matrix_vs_oracle(order, n, jitter, rotate):
    degree = 2 * order + 1
    pen = order * (order + 1) * n
    m = _rect(n, jitter=jitter, seed=order, rotate=rotate)
    U, V, a, _ = _sipg(m, order, degree, pen)
    print(f"SIPG Q{order} n={n} jitter={jitter} rotate={rotate}", flush=True)
    # the oracle numbers discontinuous dofs cell by cell, as the DofHandler does
    assert np.array_equal(_dofs(U), np.arange(m.ncells * (order + 1) ** 2))
    A = B.assemble_bilinear(a, U, V)
    R, _ = O.sipg_quad(_om(m), order, degree, pen)
    print(f"  rel err {_check_matrix(A, R):.2e}")


def _dofs(U):
    import ctypes as C
    from bcubehip import _lib as L
    nl = C.c_int32()
    nd = C.c_int64()
    L.check(L.lib().bh_space_info(U.handle, C.byref(nd), C.byref(nl)))
    out = np.zeros(U.mesh.ncells * nl.value, dtype=np.int32)
    L.check(L.lib().bh_space_dofs(U.handle, out.ctypes.data))
    return out


@pytest.mark.parametrize("terms", [("int",), ("bnd",), ("vol", "bnd"), ("vol", "int")])
def test_sipg_terms_separately(terms):
    order, n = 2, 4
    m = _rect(n, jitter=0.1, seed=7, rotate=True)
    U, V, a, _ = _sipg(m, order, 5, 12.0, terms)
    A = B.assemble_bilinear(a, U, V)
    R, _ = O.sipg_quad(_om(m), order, 5, 12.0, terms=terms)
    print(f"terms {terms}: rel err {_check_matrix(A, R):.2e}")


@pytest.mark.parametrize("order,n,jitter", [(1, 4, 0.0), (3, 4, 0.0), (2, 5, 0.2)])
def test_sipg_linear_vs_oracle(order, n, jitter):
    degree = 2 * order + 1
    pen = order * (order + 1) * n
    m = _rect(n, jitter=jitter, seed=1, rotate=jitter > 0)
    U, V, _, l = _sipg(m, order, degree, pen)
    b = B.assemble_linear(l, V)
    _, rb = O.sipg_quad(_om(m), order, degree, pen, f=_f, g=_ua)
    err = np.abs(b - rb).max() / np.abs(rb).max()
    print(f"linear Q{order} n={n}: rel err {err:.2e}")
    assert err <= RTOL


def _dg_points(m, order):
    t = np.linspace(-1.0, 1.0, order + 1)
    xi = np.array([[t[i], t[j]] for j in range(order + 1) for i in range(order + 1)])
    N = 0.25 * np.stack([(1 - xi[:, 0]) * (1 - xi[:, 1]), (1 + xi[:, 0]) * (1 - xi[:, 1]),
                         (1 + xi[:, 0]) * (1 + xi[:, 1]), (1 - xi[:, 0]) * (1 + xi[:, 1])], 1)
    X = m.coords.reshape(-1, 2)[m.c2n.reshape(-1, 4)]        # (ncells, 4, 2)
    return np.einsum("qk,ckd->cqd", N, X).reshape(-1, 2)


@pytest.mark.parametrize("jitter", [0.0, 0.2])
def test_sipg_benchmark_problem_exact(jitter):
    """bench_assemble.jl:33-96 with n = 4, Q3, degree 7, γ = 12, h = 1/4: the
    cubic solution lies in the space and the discrete system reproduces it."""
    order, n = 3, 4
    pen = order * (order + 1) / (1.0 / n)
    m = _rect(n, jitter=jitter, seed=5, rotate=jitter > 0)
    U, V, a, l = _sipg(m, order, 2 * order + 1, pen)
    A = B.assemble_bilinear(a, U, V)
    b = B.assemble_linear(l, V)
    x = spla.spsolve(A.tocsc(), b)
    err = np.abs(x - _ua(_dg_points(m, order))).max()
    print(f"benchmark problem jitter={jitter}: max nodal error {err:.2e}")
    assert err < 1e-10


def test_sipg_device_assembly_bitwise_repeatable():
    import torch
    order, n = 3, 6
    m = _rect(n, jitter=0.1, seed=2, rotate=True)
    U, V, a, _ = _sipg(m, order, 7, 72.0)
    plan = B.FaceBilinearPlan(a, U, V)
    host = plan.assemble_host()
    d = torch.empty(plan.nnz, dtype=torch.float64, device="cuda")
    for _ in range(2):
        d.fill_(np.nan)
        torch.cuda.synchronize()
        plan.assemble_device(d.data_ptr())
        B.lib().bh_sync(m.backend.ctx)
        assert np.array_equal(d.cpu().numpy(), host)
    assert np.array_equal(plan.assemble_host(), host)


def test_sipg_boundary_linear_needs_discontinuous_space():
    m = _rect(3)
    U = B.TrialFESpace(B.FunctionSpace("Lagrange", 1), m)
    V = B.TestFESpace(U)
    dΓb = B.Measure(B.BoundaryFaceDomain(m), 3)
    with pytest.raises(B.BcubeHipError):
        B.assemble_linear(lambda v: B.integrate(B.side_n(v) * 1.0, dΓb), V)


def test_sipg_direct_path_equals_gather_path():
    """Cell-major DG numbering takes the direct path (off-diagonal face blocks
    stored in place, diagonal blocks summed per cell); a permuted numbering
    takes the virtual-cell row gather.  Same sums in the same order: the two
    matrices agree entry for entry after the permutation."""
    order, n = 3, 5
    m = _rect(n, jitter=0.15, seed=4, rotate=True)
    U, V, a, _ = _sipg(m, order, 7, 60.0)
    A = B.assemble_bilinear(a, U, V)
    U2, V2, a2, _ = _sipg(m, order, 7, 60.0)
    nd = U2.get_ndofs()
    perm = np.random.default_rng(0).permutation(nd)
    U2._scalar.iglob = (perm[U2._scalar.iglob - 1] + 1).astype(np.int64)
    assert not np.array_equal(_dofs(U2), np.arange(nd))
    A2 = B.assemble_bilinear(a2, U2, V2)
    back = A2[perm][:, perm].tocsc()
    back.sort_indices()
    assert np.array_equal(back.indptr, A.indptr) and np.array_equal(back.indices, A.indices)
    assert np.array_equal(back.data, A.data)


@pytest.mark.parametrize("order", [1, 2])
def test_sipg_forms_on_a_continuous_space(order):
    """Face terms over a continuous space go through the row gather.  Its
    matrix is Pᵀ A_dg P with P the injection of continuous dofs into the
    cell-by-cell discontinuous ones (each cell's local basis is the same)."""
    import scipy.sparse as sp
    n = 4
    m = _rect(n, jitter=0.1, seed=8, rotate=True)
    Uc = B.TrialFESpace(B.FunctionSpace("Lagrange", order), m)
    Vc = B.TestFESpace(Uc)
    _, _, a, _ = _sipg(m, order, 2 * order + 1, 10.0)
    a_c = lambda u, v: a(u, v)  # noqa: E731  (the same forms, placeholders of Uc)
    A = B.assemble_bilinear(a_c, Uc, Vc)
    R, _ = O.sipg_quad(_om(m), order, 2 * order + 1, 10.0)
    dofs = _dofs(Uc)
    nl = (order + 1) ** 2
    P = sp.csr_matrix((np.ones(dofs.size), (np.arange(dofs.size), dofs)),
                      shape=(m.ncells * nl, Uc.get_ndofs()))
    ref = (P.T @ R @ P).toarray()
    err = np.abs(A.toarray() - ref).max() / np.abs(ref).max()
    print(f"continuous Q{order}: rel err {err:.2e}")
    assert err <= RTOL


def test_affine_fe_system_solves_the_benchmark_problem():
    """AffineFESystem(a, l, U, V) and solve (affine_fe_system.jl:45-119) on the
    Poisson DG benchmark: no Dirichlet dofs (Nitsche terms), the cubic
    solution reproduced."""
    order, n = 3, 4
    m = _rect(n)
    U, V, a, l = _sipg(m, order, 7, order * (order + 1) * n)
    sol = B.AffineFESystem(a, l, U, V).solve()
    err = np.abs(sol.get_dof_values() - _ua(_dg_points(m, order))).max()
    assert err < 1e-10


def test_affine_fe_system_with_dirichlet_lifting():
    """The reference's 3D heat solve (test_assembler.jl:457-509) through
    AffineFESystem: T = 260 on xmin, the cell-centre values of
    260 + (q/η) s (1 - s/2) within 2e-14."""
    m = B.hexa_mesh(5, 5, 5, xmin=-0.5, xmax=0.5, ymin=-0.5, ymax=0.5, zmin=-0.5, zmax=0.5)
    lam, q = 100.0, 1500.0
    U = B.TrialFESpace(B.FunctionSpace("Lagrange", 2), m, {"xmin": 260.0})
    V = B.TestFESpace(U)
    dΩ = B.Measure(B.CellDomain(m), 5)
    sys_ = B.AffineFESystem(lambda u, v: B.integrate(lam * B.dot(B.grad(u), B.grad(v)), dΩ),
                            lambda v: B.integrate(q * v, dΩ), U, V)
    x = sys_.solve().get_dof_values()
    centre = [U.get_dofs(c)[13] - 1 for c in range(m.ncells)]
    xc = np.array([m.nodes()[m.cell_nodes(c)].mean(axis=0) for c in range(m.ncells)])
    s = xc[:, 0] + 0.5
    Tca = 260.0 + (q / lam) * s * (1.0 - 0.5 * s)
    assert np.abs(x[centre] - Tca).max() / np.abs(Tca).max() < 2.0e-14
