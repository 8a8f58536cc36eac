"""Pins the oracle's assembly with distinct trial and test spaces and its
linear forms with an FE-function coefficient.

  * The reference's own known answers: P2-P1 on one triangle
    (test_assembler.jl:241-269: ∫(u v) with u = FEFunction(P2, 1:6) tested on
    P1 gives 97/120, 4/5, 107/120) and the discontinuous Q1 transport form on
    one 4 x 6 quad (test_assembler.jl:40-68: ∫((c u)⋅∇v) with u = α).
  * Nesting: Q1 ⊂ Q2 and P1 ⊂ P2 on the same cells, so with P[k][j] the
    closed-form order-1 basis function j at the order-2 Lagrange node k
    (reference nodes typed in here, not taken from the oracle),
    A(U=1, V=2) = A22 P and A(U=2, V=1) = Pᵀ A22, with A22 the square order-2
    matrix (pinned by the reference fixtures elsewhere), on jittered meshes,
    for mass and stiffness.  A transposed block, a wrong basis or a wrong dof
    map fails it.
  * ∫(u v) with u in the order-1 space equals M(U=1, V) u, and constants: Σ of
    the rectangular mass is the area, the rectangular stiffness kills 1.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import oracle as O

# order-2 reference nodes in the reference's Lagrange order (x fastest on the
# square and cube, hypercube.jl:56-92; vertices then edge midpoints on the
# triangle, triangle.jl:4-64)
_T = np.array([-1.0, 0.0, 1.0])
NODES2 = {
    "Quad4": np.array([[x, y] for y in _T for x in _T]),
    "Tri3": np.array([[0, 0], [1, 0], [0, 1], [0.5, 0], [0.5, 0.5], [0, 0.5]], float),
    "Hexa8": np.array([[x, y, z] for z in _T for y in _T for x in _T]),
}


def _p1(kind, xi):
    """Closed-form order-1 basis at reference point xi."""
    if kind == "Tri3":
        x, y = xi
        return np.array([1 - x - y, x, y])
    lo = [(1 - t) / 2 for t in xi]
    hi = [(1 + t) / 2 for t in xi]
    if kind == "Quad4":
        return np.array([lo[0] * lo[1], hi[0] * lo[1], lo[0] * hi[1], hi[0] * hi[1]])
    return np.array([a * b * c for c in (lo[2], hi[2]) for b in (lo[1], hi[1])
                     for a in (lo[0], hi[0])])


def _mesh(kind, n):
    if kind == "Hexa8":
        o = O.hexa_mesh(n + 1, n + 1, n + 1)
        d = 3
    else:
        o = O.rectangle_mesh(n + 1, n + 1)
        d = 2
        if kind == "Tri3":
            X = o.coords.reshape(-1, 2)
            q = o.c2n.reshape(-1, 4)
            tris = np.concatenate([q[:, [0, 1, 2]], q[:, [0, 2, 3]]], axis=1).reshape(-1, 3)
            o = O.OMesh.uniform(2, X, O.CT["Tri3"], tris)
    X = o.coords.reshape(-1, d).copy()
    inner = np.all((X > 1e-12) & (X < 1 - 1e-12), axis=1)
    X[inner] += np.random.default_rng(3).uniform(-0.15, 0.15, (inner.sum(), d)) / n
    o.coords = np.ascontiguousarray(X.reshape(-1))
    return o


def _prolongation(o, kind):
    ig1, off1, n1 = O.dofhandler(o, 1)
    ig2, off2, n2 = O.dofhandler(o, 2)
    nl1, nl2 = len(_p1(kind, NODES2[kind][0])), len(NODES2[kind])
    P = np.zeros((n2, n1))
    for c in range(o.ncells):
        d1 = ig1[off1[c]:off1[c] + nl1] - 1
        d2 = ig2[off2[c]:off2[c] + nl2] - 1
        for k in range(nl2):
            P[d2[k], d1] = _p1(kind, NODES2[kind][k])
    return P


def _dense(res):
    cp, rv, nz, shape = res
    return sp.csc_matrix((nz, rv - 1, cp - 1), shape=shape).toarray()


@pytest.mark.parametrize("kind,n,deg", [("Quad4", 3, 4), ("Tri3", 3, 4), ("Hexa8", 2, 4)])
@pytest.mark.parametrize("term", ["mass", "stiffness"])
def test_rect_equals_nested_square(kind, n, deg, term):
    o = _mesh(kind, n)
    eta = np.random.default_rng(9).uniform(0.5, 2.0, o.ncells)
    e = eta if term == "stiffness" else None
    cp, rv, nz, n22 = O.assemble_bilinear(o, 2, [term], eta=e, degree=deg)
    A22 = sp.csc_matrix((nz, rv - 1, cp - 1), shape=(n22, n22)).toarray()
    P = _prolongation(o, kind)
    A12 = _dense(O.assemble_bilinear_rect(o, 1, 2, [term], eta=e, degree=deg))
    A21 = _dense(O.assemble_bilinear_rect(o, 2, 1, [term], eta=e, degree=deg))
    s = np.abs(A22).max()
    assert A12.shape == (n22, P.shape[1]) and A21.shape == (P.shape[1], n22)
    assert np.abs(A12 - A22 @ P).max() <= 1e-13 * s
    assert np.abs(A21 - P.T @ A22).max() <= 1e-13 * s
    if term == "mass":
        assert abs(A12.sum() - 1.0) <= 1e-13
    else:
        assert np.abs(A12.sum(axis=1)).max() <= 1e-12 * s
        assert np.abs(A21.sum(axis=0)).max() <= 1e-12 * s


def test_p2_p1_known_answer():
    """test_assembler.jl:241-269."""
    o = O.OMesh.uniform(2, np.array([[0, 0], [1, 0], [0, 1]], float), O.CT["Tri3"], [[0, 1, 2]])
    b = O.assemble_linear_fe(o, 2, np.arange(1.0, 7.0), 1, "uv", degree=3)
    assert np.allclose(b, [97 / 120, 4 / 5, 107 / 120], rtol=0, atol=1e-15)


def test_q1_transport_known_answer():
    """test_assembler.jl:40-68: one quad [-2,2] x [-3,3], discontinuous Q1,
    u = α everywhere, c = (0.5, 1)."""
    alpha, c, dx, dy = 10.0, np.array([0.5, 1.0]), 4.0, 6.0
    X = np.array([[-1, -1], [1, -1], [1, 1], [-1, 1]], float) * [2.0, 3.0]
    o = O.OMesh.uniform(2, X, O.CT["Quad4"], [[0, 1, 2, 3]])
    b = O.assemble_linear_fe(o, 1, np.full(4, alpha), 1, "cu_grad", c=c, degree=3,
                             continuous=False)
    ref = [-alpha / 2 * (c[0] * dy + c[1] * dx), alpha / 2 * (c[0] * dy - c[1] * dx),
           alpha / 2 * (-c[0] * dy + c[1] * dx), alpha / 2 * (c[0] * dy + c[1] * dx)]
    assert np.allclose(b, ref, rtol=1e-14, atol=0)


@pytest.mark.parametrize("kind", ["Quad4", "Tri3"])
def test_fe_linear_form_is_rect_mass_times_u(kind):
    o = _mesh(kind, 3)
    u = np.random.default_rng(1).normal(size=O.dofhandler(o, 1)[2])
    b = O.assemble_linear_fe(o, 1, u, 2, "uv", degree=4)
    M = _dense(O.assemble_bilinear_rect(o, 1, 2, ["mass"], degree=4))
    assert np.abs(b - M @ u).max() <= 1e-14 * np.abs(b).max()
