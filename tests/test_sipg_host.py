"""The symmetric interior-penalty (Poisson DG) restatement of the oracle and
the symbolic lowering of face integrands, on the CPU.

The oracle is pinned by what the mathematics fixes, so that a dropped term, a
wrong sign, a wrong normal or a wrong side permutation fails one check:
  - consistency: the benchmark problem of benchmark/bench_assemble.jl:33-96
    (u = 3x + y² + 2x³ in Q3, Dirichlet data by Nitsche terms) is solved
    exactly by the discrete system;
  - symmetry of the assembled matrix;
  - constants are in the kernel of a_Ω + a_Γ (no boundary terms);
  - the penalty part of a_Γ, (γ/h)·Σ_f ∫ jump(v,n)⋅jump(u,n), evaluated on the
    indicator of one cell, is (γ/h) times the length of that cell's interior
    edges.
"""
import numpy as np
import pytest

from oracle import oracle as O


def _rect(n, lo=-0.5, hi=0.5, jitter=0.0, seed=0, rotate=False):
    """Structured Quad4 mesh of n x n cells (node (iy*(n+1) + ix), cells x-fastest),
    optional interior-node jitter and cyclic rotation of odd cells' nodes."""
    t = np.linspace(lo, hi, n + 1)
    X = np.array([[x, y] for y in t for x in t])
    if jitter:
        rng = np.random.default_rng(seed)
        inner = (np.abs(X[:, 0] - lo) > 1e-12) & (np.abs(X[:, 0] - hi) > 1e-12) & \
                (np.abs(X[:, 1] - lo) > 1e-12) & (np.abs(X[:, 1] - hi) > 1e-12)
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
    c2n = np.array(cells, dtype=np.int64).reshape(-1)
    return O.OMesh(2, X.reshape(-1), np.zeros(n * n, dtype=np.int32),
                   np.arange(n * n + 1, dtype=np.int64) * 4, c2n)


def _u(x):
    return 3 * x[0] + x[1] ** 2 + 2 * x[0] ** 3


def _f(x):
    return -2.0 - 12.0 * x[0]


@pytest.mark.parametrize("jitter,rotate", [(0.0, False), (0.15, True)])
def test_oracle_sipg_solves_cubic_exactly(jitter, rotate):
    n, order = 4, 3
    degree = 2 * order + 1
    gamma = order * (order + 1)
    m = _rect(n, jitter=jitter, rotate=rotate)
    A, b = O.sipg_quad(m, order, degree, gamma / (1.0 / n), f=_f, g=_u)
    import scipy.sparse.linalg as spla
    x = spla.spsolve(A.tocsc(), b)
    ua = np.array([_u(p) for p in O.dg_node_coords(m, order)])
    assert np.abs(x - ua).max() < 1e-10
    assert abs(A - A.T).max() <= 1e-13 * abs(A).max()


def test_oracle_sipg_constants_in_kernel_and_penalty_measure():
    n, order, degree = 3, 2, 5
    m = _rect(n, jitter=0.2, seed=3, rotate=True)
    nl = (order + 1) ** 2
    A, _ = O.sipg_quad(m, order, degree, 7.0, terms=("vol", "int"))
    assert np.abs(A @ np.ones(A.shape[0])).max() < 1e-12 * abs(A).max()
    A1, _ = O.sipg_quad(m, order, degree, 1.0, terms=("int",))
    A0, _ = O.sipg_quad(m, order, degree, 0.0, terms=("int",))
    P = A1 - A0
    X = m.coords.reshape(-1, 2)
    for c in range(n * n):
        ind = np.zeros(A.shape[0])
        ind[c * nl:(c + 1) * nl] = 1.0
        nodes = m.c2n[4 * c:4 * c + 4]
        length = 0.0
        for s in range(4):
            a, bb = nodes[s], nodes[(s + 1) % 4]
            edge = {int(a), int(bb)}
            shared = sum(1 for d in range(n * n) if d != c and
                         edge <= set(int(v) for v in m.c2n[4 * d:4 * d + 4]))
            if shared:
                length += np.linalg.norm(X[a] - X[bb])
        assert abs(ind @ (P @ ind) - length) < 1e-12


# ---- symbolic lowering of face integrands -----------------------------------------
def _forms():
    import bcubehip as B
    from bcubehip.forms import Placeholder

    class _Mesh:
        pass

    mesh = _Mesh()
    dG = B.Measure(B.InteriorFaceDomain(mesh), 7)
    dGb = B.Measure(B.BoundaryFaceDomain(mesh), 7)
    return B, Placeholder, B.get_face_normals(dG), B.get_face_normals(dGb)


def test_lower_sipg_interior_and_boundary():
    B, P = _forms()[:2]
    _, _, n, nb = _forms()
    from bcubehip.forms import lower_face_bilinear, lower_face_linear
    u, v = P("trial", None), P("test", None)
    pen = 48.0
    e = (-B.dot(B.jump(v, n), B.avg(B.grad(u))) - B.dot(B.avg(B.grad(v)), B.jump(u, n))
         + pen * B.dot(B.jump(v, n), B.jump(u, n)))
    c = lower_face_bilinear(e, False).reshape(4, 4)
    # (t, r) rows; columns A, B, C, E of the kernels' monomials
    assert np.array_equal(c[:, 0], np.zeros(4))
    assert np.array_equal(c[:, 1], -0.5 * np.ones(4))
    assert np.array_equal(c[:, 2], -0.5 * np.ones(4))
    assert np.array_equal(c[:, 3], [pen, -pen, -pen, pen])
    fa = lambda u, gu, v, gv, n: -v * B.dot(gu, n) - B.dot(gv, n) * u + pen * v * u  # noqa: E731
    cb = lower_face_bilinear(B.compose(fa, *map(B.side_n, (u, B.grad(u), v, B.grad(v), nb))),
                             True)
    assert np.array_equal(cb, [0.0, -1.0, -1.0, pen])
    g = B.PhysicalFunction(lambda x: x[:, 0])
    fl = lambda v, gv, n, g: -B.dot(gv, n) * g + pen * v * g  # noqa: E731
    groups = lower_face_linear(B.compose(fl, *map(B.side_n, (v, B.grad(v), nb, g))))
    assert len(groups) == 1 and groups[0][:2] == (pen, -1.0) and groups[0][2] is g


def test_lower_face_normals_of_the_other_side():
    """n⁺ = -n⁻ on a straight face: writing a term with the other side's normal
    flips its coefficient, and jump(v) = v⁻ - v⁺ expands like jump(v, n)⋅n⁻."""
    B, P, n, _ = _forms()
    from bcubehip.forms import lower_face_bilinear
    u, v = P("trial", None), P("test", None)
    a = lower_face_bilinear(B.side_n(v) * B.dot(B.side_p(B.grad(u)), B.side_p(n)), False)
    b = lower_face_bilinear(B.side_n(v) * B.dot(B.side_p(B.grad(u)), B.side_n(n)), False)
    assert np.array_equal(a, -b)
    assert a[4 * 1 + 1] == -1.0   # (t, r) = (0, 1): B = σ_0 · (-1)
    j1 = lower_face_bilinear(B.jump(v) * B.jump(u), False)
    j2 = lower_face_bilinear(B.dot(B.jump(v, n), B.side_n(n)) * B.dot(B.jump(u, n), B.side_n(n)),
                             False)
    assert np.array_equal(j1, j2)


def test_lower_face_rejects_unsupported():
    B, P, n, _ = _forms()
    from bcubehip.forms import lower_face_bilinear
    u, v = P("trial", None), P("test", None)
    with pytest.raises(NotImplementedError):
        lower_face_bilinear(B.dot(B.avg(B.grad(u)), B.avg(B.grad(v))), False)
    with pytest.raises(NotImplementedError):
        lower_face_bilinear(u * v, False)   # no side on an interior face
    with pytest.raises(NotImplementedError):
        lower_face_bilinear(B.side_n(v) * B.side_n(n), False)   # vector-valued
