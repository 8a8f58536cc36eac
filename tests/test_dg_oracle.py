"""Pins the oracle's DG transport residual (volume + upwind interior faces).

The reference cannot integrate over Hexa27 cells (no `shape` method for
Hexa27_t, src/mesh/shape.jl:18-36), so this path is pinned by mathematics:

* the triquadratic map reproduces any triquadratic field and its derivatives;
* with u = 1 the residual of every cell away from the boundary vanishes on the
  curved mesh (divergence theorem, integrated exactly by the 4-point rule for
  Q2/Q3) -- this fails for a wrong side-p permutation or a flipped normal;
* a linear u on an affine mesh gives r_i = -(c.a) h^3 w_i1 w_i2 w_i3;
* one cell with u = 1: the cell receives its inflow-face integrals and its
  downstream neighbours the matching outflow ones (upwind selection, jump sign);
* a Hexa27 mesh placed on trilinear cells gives the Hexa8 residual.
"""
import numpy as np
import pytest

import bcubehip as bh
from oracle import oracle as O

C = np.array([1.0, 0.5, 0.25])
HOST = bh.BcubeBackendHostOnly()
W1 = {1: [0.5, 0.5], 2: [1 / 6, 4 / 6, 1 / 6], 3: [1 / 8, 3 / 8, 3 / 8, 1 / 8]}


def _om(m):
    return O.OMesh(m.dim, m.coords, m.celltypes, m.c2n_ptr, m.c2n)


def _interior(n):
    cid = np.arange(n ** 3)
    k = np.stack([cid % n, (cid // n) % n, cid // (n * n)])
    return np.all((k > 0) & (k < n - 1), axis=0)


def _idx(N):
    return np.array([[i1, i2, i3] for i3 in range(N) for i2 in range(N) for i1 in range(N)])


def test_hexa27_map_reproduces_triquadratic_field():
    rng = np.random.default_rng(0)
    coef = rng.standard_normal((3, 3, 3, 3))  # component, powers of xi, eta, zeta

    def P(x):
        return np.array([sum(coef[d, a, b, e] * x[0] ** a * x[1] ** b * x[2] ** e
                             for a in range(3) for b in range(3) for e in range(3))
                         for d in range(3)])

    def dP(x):
        J = np.zeros((3, 3))
        for d in range(3):
            for a in range(3):
                for b in range(3):
                    for e in range(3):
                        c = coef[d, a, b, e]
                        if a:
                            J[d, 0] += c * a * x[0] ** (a - 1) * x[1] ** b * x[2] ** e
                        if b:
                            J[d, 1] += c * b * x[0] ** a * x[1] ** (b - 1) * x[2] ** e
                        if e:
                            J[d, 2] += c * e * x[0] ** a * x[1] ** b * x[2] ** (e - 1)
        return J

    ref = bh.mesh._H27_REF.astype(float)
    nodes = np.array([P(r) for r in ref])
    for _ in range(5):
        xi = rng.uniform(-1, 1, 3)
        x, J, Ji, det = O.geometry27(nodes, xi)
        np.testing.assert_allclose(x, P(xi), rtol=0, atol=1e-12)
        np.testing.assert_allclose(J, dP(xi), rtol=0, atol=1e-11)
        np.testing.assert_allclose(Ji @ J, np.eye(3), atol=1e-10)
        assert det == pytest.approx(abs(np.linalg.det(dP(xi))), rel=1e-11)


@pytest.mark.parametrize("p", [2, 3])
def test_constant_field_interior_residual_vanishes_on_curved_mesh(p):
    n, N = 4, p + 1
    m = bh.hexa27_mesh(n + 1, n + 1, n + 1, backend=HOST)
    bh.sinusoidal_map(m, 0.05, seed=0)
    u = np.ones((m.ncells * N ** 3, 2))
    r = O.dg_transport(_om(m), p, C, u, 2 * N - 1)
    rc = r.reshape(m.ncells, N ** 3, 2)
    scale = np.abs(r).max()
    assert scale > 1e-3
    assert np.abs(rc[_interior(n)]).max() < 1e-14 * max(scale, 1.0)
    # sum of test functions is 1: the residual sums to zero
    assert np.abs(r.sum(axis=0)).max() < 1e-14


@pytest.mark.parametrize("p", [1, 2, 3])
def test_linear_field_affine_closed_form(p):
    n, N = 4, p + 1
    h = 1.0 / n
    m = bh.hexa27_mesh(n + 1, n + 1, n + 1, backend=HOST)
    X = m.nodes()
    c2n = m.c2n.reshape(-1, 27)
    idx = _idx(N)
    pts = -1 + 2 * idx / p
    xd = X[c2n[:, 0]][:, None, :] + (pts[None] + 1) / 2 * h
    a, b0 = np.array([0.3, -0.7, 1.1]), 0.2
    r = O.dg_transport(_om(m), p, C, (xd @ a + b0).reshape(-1, 1), 2 * N - 1)
    w = np.array(W1[p])
    expect = -(C @ a) * h ** 3 * w[idx[:, 0]] * w[idx[:, 1]] * w[idx[:, 2]]
    got = r.reshape(-1, N ** 3)[_interior(n)]
    np.testing.assert_allclose(got, np.broadcast_to(expect, got.shape), rtol=0, atol=1e-15)


@pytest.mark.parametrize("p", [1, 2, 3])
def test_upwind_single_cell(p):
    n, N = 4, p + 1
    h = 1.0 / n
    m = bh.hexa27_mesh(n + 1, n + 1, n + 1, backend=HOST)
    K = int(np.flatnonzero(_interior(n))[0])
    u = np.zeros((m.ncells, N ** 3))
    u[K] = 1.0
    r = O.dg_transport(_om(m), p, C, u.reshape(-1, 1), 2 * N - 1).reshape(-1, N ** 3)
    w = np.array(W1[p])
    idx = _idx(N)
    want = np.zeros_like(r)
    for d, step in ((0, 1), (1, n), (2, n * n)):
        o1, o2 = [e for e in range(3) if e != d]
        fw = w[idx[:, o1]] * w[idx[:, o2]] * h * h
        on_minus = idx[:, d] == 0
        want[K] += np.where(on_minus, -C[d] * fw, 0.0)         # inflow face of K
        want[K + step] += np.where(on_minus, C[d] * fw, 0.0)   # K's outflow, seen downstream
    np.testing.assert_allclose(r, want, rtol=0, atol=1e-15)


@pytest.mark.parametrize("p", [1, 3])
def test_hexa27_on_trilinear_cells_matches_hexa8(p):
    n, N = 4, p + 1
    m8 = bh.hexa_mesh(n + 1, n + 1, n + 1, backend=HOST)
    bh.jitter_nodes(m8, 0.15, 1.0 / n, seed=4)
    m27 = bh.hexa8_to_hexa27(m8, backend=HOST)
    u = np.random.default_rng(1).standard_normal((m8.ncells * N ** 3, 3))
    r8 = O.dg_transport(_om(m8), p, C, u, 2 * N - 1)
    r27 = O.dg_transport(_om(m27), p, C, u, 2 * N - 1)
    assert np.abs(r8 - r27).max() < 1e-13 * np.abs(r8).max()


def test_hexa27_mesh_nodes_on_lattice():
    m = bh.hexa27_mesh(3, 2, 2, backend=HOST)
    X = m.nodes()
    c2n = m.c2n.reshape(-1, 27)
    ref = bh.mesh._H27_REF
    for c in range(m.ncells):
        centre = X[c2n[c, 26]]
        half = np.array([0.25, 0.5, 0.5])
        np.testing.assert_allclose(X[c2n[c]], centre + ref * half, atol=1e-15)
