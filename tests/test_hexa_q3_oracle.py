"""Pins the oracle's continuous Q3 hexahedra (64 dofs per cell, 4^3 Gauss points).

The reference has no test of its own for this element, so the pins are
mathematical:
  * the assembled matrix equals Σ_c P_cᵀ K_c P_c, with K_c built here from 1D
    Lagrange matrices (numpy's Gauss-Legendre and polynomial fits, tensor
    products in the reference's x-fastest dof order, hypercube.jl:56-92) on
    cells that are axis-aligned boxes of different sizes, and P_c the oracle's
    numbering: a wrong basis order, gradient, weight or Jacobian fails it;
  * constants: K·1 = 0 and 1ᵀM1 is the volume on a jittered mesh;
  * the numbering (dofhandler.jl:37-156 restated): vertex, face and interior
    dofs of a global id sit at one physical point in every cell that has it;
    where two cells disagree it is an edge dof, placed at the mirror image
    along its edge (the reference copies edge dofs of 3D cells through the
    face-matching branch, dofhandler.jl:239-290, entity.jl:487-502).  Which
    edges flip follows the restated loop: no reference fixture pins that for
    Q3 hexahedra (parity unpinned beyond the mirror property).
"""
import numpy as np
import scipy.sparse as sp

from oracle import oracle as O

XI = np.array([-1.0, -1.0 / 3.0, 1.0 / 3.0, 1.0])


def _lagrange_1d():
    """1D cubic Lagrange polynomials on XI, as numpy polynomials."""
    P = []
    for i in range(4):
        y = np.zeros(4)
        y[i] = 1.0
        P.append(np.polynomial.Polynomial.fit(XI, y, 3, domain=[-1, 1], window=[-1, 1]))
    return P


def _mats_1d(h):
    P = _lagrange_1d()
    x, w = np.polynomial.legendre.leggauss(6)
    v = np.array([p(x) for p in P])
    d = np.array([p.deriv()(x) for p in P])
    M = (v * w) @ v.T * (h / 2)
    K = (d * w) @ d.T * (2 / h)
    return M, K


def _box_mesh(n):
    o = O.hexa_mesh(n + 1, n + 1, n + 1)
    X = o.coords.reshape(-1, 3).copy()
    # each axis mapped on its own: the cells stay axis-aligned boxes
    X[:, 0] = X[:, 0] + 0.15 * np.sin(np.pi * X[:, 0])
    X[:, 1] = X[:, 1] ** 1.3
    X[:, 2] = 2.0 * X[:, 2] + 0.1 * X[:, 2] ** 2
    o.coords = np.ascontiguousarray(X.reshape(-1))
    return o


def _csc(res):
    cp, rv, nz, nd = res
    return sp.csc_matrix((nz, rv - 1, cp - 1), shape=(nd, nd))


def test_q3_hexa_matrix_equals_tensor_element_sum():
    n = 3
    o = _box_mesh(n)
    eta = np.random.default_rng(4).uniform(0.5, 2.0, o.ncells)
    A = _csc(O.assemble_bilinear(o, 3, ["stiffness", "mass"], eta=eta, degree=6)).toarray()
    ig, off, nd = O.dofhandler(o, 3)
    ig = ig.reshape(-1, 64) - 1
    X = o.coords.reshape(-1, 3)
    c2n = o.c2n.reshape(-1, 8)
    R = np.zeros((nd, nd))
    for c in range(o.ncells):
        h = X[c2n[c]].max(0) - X[c2n[c]].min(0)
        (Mx, Kx), (My, Ky), (Mz, Kz) = (_mats_1d(hh) for hh in h)
        # dof k = i + 4 j + 16 l: x fastest, so kron(z, kron(y, x))
        K = np.kron(Mz, np.kron(My, Kx)) + np.kron(Mz, np.kron(Ky, Mx)) + np.kron(Kz, np.kron(My, Mx))
        M = np.kron(Mz, np.kron(My, Mx))
        g = ig[c]
        R[np.ix_(g, g)] += eta[c] * K + M
    assert np.abs(A - R).max() <= 1e-13 * np.abs(R).max()


def test_q3_hexa_constants_on_jittered_mesh():
    n = 3
    o = O.hexa_mesh(n + 1, n + 1, n + 1)
    X = o.coords.reshape(-1, 3).copy()
    rng = np.random.default_rng(0)
    inner = np.all((X > 1e-12) & (X < 1 - 1e-12), axis=1)
    X[inner] += rng.uniform(-0.15, 0.15, (inner.sum(), 3)) / n
    o.coords = np.ascontiguousarray(X.reshape(-1))
    K = _csc(O.assemble_bilinear(o, 3, ["stiffness"], degree=6))
    M = _csc(O.assemble_bilinear(o, 3, ["mass"], degree=6))
    one = np.ones(K.shape[0])
    assert np.abs(K @ one).max() <= 1e-12 * abs(K).max()
    assert abs(one @ (M @ one) - 1.0) <= 1e-13
    assert abs(K - K.T).max() == 0.0


def test_q3_hexa_numbering_positions():
    n = 3
    o = _box_mesh(n)
    ig, off, nd = O.dofhandler(o, 3)
    ig = ig.reshape(-1, 64) - 1
    X = o.coords.reshape(-1, 3)
    c2n = o.c2n.reshape(-1, 8)
    first = {}
    mirrored = 0
    for c in range(o.ncells):
        lo, hi = X[c2n[c]].min(0), X[c2n[c]].max(0)
        for k in range(64):
            t = (k % 4, (k // 4) % 4, k // 16)
            p = lo + (XI[list(t)] + 1) / 2 * (hi - lo)
            g = int(ig[c, k])
            if g not in first:
                first[g] = p
                continue
            d = np.abs(first[g] - p) > 1e-12
            if not d.any():
                continue
            # an edge dof (two local coordinates on the cell's boundary), at
            # the mirror image of the first cell's point along the edge
            on_bnd = sum(1 for a in t if a in (0, 3))
            assert on_bnd == 2 and d.sum() == 1, (c, k)
            ax = int(np.argmax(d))
            assert t[ax] in (1, 2)
            assert abs(first[g][ax] + p[ax] - (lo[ax] + hi[ax])) <= 1e-12
            mirrored += 1
    assert len(first) == nd == (3 * n + 1) ** 3
    assert mirrored > 0
