"""Pins the oracle's triangle path (Tri3 geometry, P1..P4 Lagrange, the
reference's inline triangle rules) and mixed triangle/quad assembly.

* rules: exact on every monomial up to their degree (∫ x^a y^b = a! b! / (a+b+2)!);
* bases: Kronecker delta at the reference's node coordinates
  (triangle.jl:120-145 for P1..P3; the P4 extension at its own lattice),
  partition of unity, gradients against central differences;
* Tri3: J, J⁻¹ and |J| of an affine map;
* mixed meshes: Σ M = area, K·1 = 0 and the patch tests K·u = -M·Δu on
  interior dofs for u = x (all orders) and u = x² + xy (P2..P4, Q2..Q4).
"""
from math import factorial

import numpy as np
import pytest

import bcubehip as B
from oracle import oracle as O

HOST = B.BcubeBackendHostOnly()


@pytest.mark.parametrize("degree,npts", [(1, 1), (2, 3), (3, 4), (4, 6), (5, 7), (6, 12),
                                         (7, 16), (8, 16)])
def test_triangle_rules_exact(degree, npts):
    w, x = O.quadrature("Triangle", degree)
    assert w.size == npts
    exact_to = 8 if degree == 7 else degree
    for a in range(exact_to + 1):
        for b in range(exact_to + 1 - a):
            got = np.sum(w * x[:, 0] ** a * x[:, 1] ** b)
            want = factorial(a) * factorial(b) / factorial(a + b + 2)
            assert got == pytest.approx(want, rel=1e-13, abs=1e-16), (a, b)


def test_femquad_rules_are_the_published_ones():
    """FEMQuad 0.3's GLTRI1/3/4 (quadrature.jl:280-288): exactness alone also
    admits other rules of these degrees (the edge-midpoint rule of degree 2,
    a 4-point rule with positive weights), so the first moment beyond each
    rule's degree tells them apart: centroid (1/3)^2/2 for x^2; interior
    points (8/27 + 2/216)/6 = 11/216 for x^3 (the edge midpoints give 1/24);
    Strang & Fix (-27/81 + 25*(0.6^4 + 2*0.2^4))/96 for x^4 with the
    negative centroid weight."""
    w, x = O.quadrature("Triangle", 1)
    assert np.sum(w * x[:, 0] ** 2) == pytest.approx(1 / 18, rel=1e-14)
    w, x = O.quadrature("Triangle", 2)
    assert np.all(x > 0) and np.sum(w * x[:, 0] ** 3) == pytest.approx(11 / 216, rel=1e-14)
    w, x = O.quadrature("Triangle", 3)
    assert w.min() < 0
    want = (-27 / 81 + 25 * (0.6 ** 4 + 2 * 0.2 ** 4)) / 96
    assert np.sum(w * x[:, 0] ** 4) == pytest.approx(want, rel=1e-14)
    with pytest.raises(RuntimeError):
        O.quadrature("Triangle", 9)


def _tri_nodes(p):
    v = np.array([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0]])
    if p == 1:
        return v
    if p == 2:
        return np.vstack([v, (v[0] + v[1]) / 2, (v[1] + v[2]) / 2, (v[2] + v[0]) / 2])
    if p == 3:
        e = lambda i, j: [(2 / 3) * v[i] + (1 / 3) * v[j], (1 / 3) * v[i] + (2 / 3) * v[j]]
        return np.vstack([v] + [e(0, 1), e(1, 2), e(2, 0)] + [[v.sum(0) / 3]])
    bary = [(4, 0, 0), (0, 4, 0), (0, 0, 4), (3, 1, 0), (2, 2, 0), (1, 3, 0), (0, 3, 1),
            (0, 2, 2), (0, 1, 3), (1, 0, 3), (2, 0, 2), (3, 0, 1), (2, 1, 1), (1, 2, 1), (1, 1, 2)]
    return np.array([(b[0] * v[0] + b[1] * v[1] + b[2] * v[2]) / 4 for b in bary])


@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_triangle_basis_delta_unity_and_gradients(p):
    nodes = _tri_nodes(p)
    for k, xi in enumerate(nodes):
        val, _ = O.shape_functions("Triangle", p, xi)
        np.testing.assert_allclose(val, np.eye(len(nodes))[k], atol=1e-14)
    rng = np.random.default_rng(p)
    for _ in range(4):
        xi = rng.dirichlet([1, 1, 1])[:2]
        val, grad = O.shape_functions("Triangle", p, xi)
        assert val.sum() == pytest.approx(1.0, abs=1e-14)
        np.testing.assert_allclose(grad.sum(axis=0), 0.0, atol=1e-12)
        h = 1e-6
        for a in range(2):
            e = np.zeros(2)
            e[a] = h
            fd = (O.shape_functions("Triangle", p, xi + e)[0]
                  - O.shape_functions("Triangle", p, xi - e)[0]) / (2 * h)
            np.testing.assert_allclose(grad[:, a], fd, atol=1e-7)


def test_tri3_geometry():
    X = np.array([[0.3, 0.1], [1.4, 0.5], [0.2, 1.7]])
    J, Ji, det, x = O.geometry("Tri3", X, np.array([0.25, 0.5]))
    np.testing.assert_allclose(J, np.column_stack([X[1] - X[0], X[2] - X[0]]), atol=1e-15)
    np.testing.assert_allclose(Ji @ J, np.eye(2), atol=1e-14)
    assert det == pytest.approx(abs(np.linalg.det(J)), rel=1e-14)
    np.testing.assert_allclose(x, 0.25 * X[0] + 0.25 * X[1] + 0.5 * X[2], atol=1e-15)


def _mixed(n, seed=0, jitter=0.2):
    m, _ = B.mixed_tri_quad_mesh(n, block=2, p_split=0.5, seed=seed, backend=HOST)
    if jitter:
        B.jitter_nodes(m, jitter, 1.0 / n, seed=seed)
    return m


def _om(m):
    return O.OMesh(m.dim, m.coords, m.celltypes, m.c2n_ptr, m.c2n)


def _dof_coords(m, p):
    """Physical position of every global dof (nodes mapped from each cell)."""
    iglob, offset, _ = O.dofhandler(_om(m), p)
    nd = int(iglob.max())
    X = np.zeros((nd, 2))
    Xn = m.nodes()
    for c in range(m.ncells):
        cn = m.cell_nodes(c)
        if len(cn) == 3:
            ref = _tri_nodes(p)
            P = Xn[cn]
            phys = P[0] + ref[:, :1] * (P[1] - P[0]) + ref[:, 1:] * (P[2] - P[0])
        else:
            t = np.linspace(-1, 1, p + 1)
            ref = np.array([[t[i], t[j]] for j in range(p + 1) for i in range(p + 1)])
            P = Xn[cn]
            s, u = ref[:, :1], ref[:, 1:]
            phys = ((1 - s) * (1 - u) * P[0] + (1 + s) * (1 - u) * P[1] + (1 + s) * (1 + u) * P[2]
                    + (1 - s) * (1 + u) * P[3]) / 4
        nl = phys.shape[0]
        X[iglob[offset[c]:offset[c] + nl] - 1] = phys
    return X


@pytest.mark.parametrize("p,deg", [(1, 4), (2, 4), (3, 6), (4, 8)])
def test_mixed_mesh_mass_and_patch_tests(p, deg):
    m = _mixed(6, seed=p, jitter=0.0 if p > 1 else 0.2)
    o = _om(m)
    K = O.csc_to_scipy(*O.assemble_bilinear(o, p, ["stiffness"], degree=deg))
    M = O.csc_to_scipy(*O.assemble_bilinear(o, p, ["mass"], degree=deg))
    n = K.shape[0]
    assert M.sum() == pytest.approx(1.0, rel=1e-13)
    assert np.abs(K @ np.ones(n)).max() < 1e-12
    X = _dof_coords(m, p)
    inner = np.all((X > 1e-9) & (X < 1 - 1e-9), axis=1)
    assert np.abs((K @ X[:, 0])[inner]).max() < 1e-12
    if p >= 2:  # u = x² + xy, -Δu = -2
        u = X[:, 0] ** 2 + X[:, 0] * X[:, 1]
        r = K @ u + 2.0 * (M @ np.ones(n))
        assert np.abs(r[inner]).max() < 1e-12
