"""The oracle's continuous spaces of per-cell order (2D Quad4/Tri3, edges at
their cells' lowest order), pinned without restating its formulas:

* one order everywhere: numbering and matrix equal the reference restatement
  (DofHandler + assemble_bilinear) bit for bit;
* continuity: the trace of every global basis function on every interior edge
  is the same from both cells.  The slot-to-node interpolation C is rebuilt
  here from the reference nodes' positions (found from the basis itself) and
  scipy's Lagrange polynomials, not from the oracle's tables;
* embedding: the space lies inside the uniform highest-order space, so its
  matrix is Pᵀ A P with A the uniform matrix (same mesh, same rule) and P the
  interpolation of each cell's functions at the uniform space's nodes;
* constants: stiffness rows sum to zero, 1ᵀM1 is the domain's area.
"""
import numpy as np
import pytest
from scipy.interpolate import lagrange

import bcubehip as B
from oracle import oracle as O

HOST = B.BcubeBackendHostOnly()
QUAD, TRI = 0, 2
VREF = {QUAD: np.array([[-1.0, -1.0], [1.0, -1.0], [1.0, 1.0], [-1.0, 1.0]]),
        TRI: np.array([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0]])}
EDGES = {QUAD: [(0, 1), (1, 2), (2, 3), (3, 0)], TRI: [(0, 1), (1, 2), (2, 0)]}
SHAPE = {QUAD: "Square", TRI: "Triangle"}


def _om(m):
    return O.OMesh(m.dim, m.coords, m.celltypes, m.c2n_ptr, m.c2n)


def _mesh(n, seed, block=2):
    m, blocks = B.mixed_tri_quad_mesh(n, block=block, p_split=0.5, seed=seed, backend=HOST)
    B.jitter_nodes(m, 0.2, 1.0 / n, seed=seed + 100)
    return m, blocks


def _orders(blocks, seed, probs=(0.4, 0.3, 0.2, 0.1)):
    nb = int(blocks.max()) + 1
    return np.random.default_rng(seed).choice([1, 2, 3, 4], size=nb, p=probs)[blocks].astype(np.int32)


_NODES = {}


def _nodes(ct, p):
    """Reference positions of the order-p Lagrange dofs, found from the basis:
    the lattice point where function a is 1 and every other is 0."""
    if (ct, p) in _NODES:
        return _NODES[(ct, p)]
    if ct == QUAD:
        g = -1.0 + 2.0 * np.arange(p + 1) / p
        cand = np.array([[x, y] for x in g for y in g])
    else:
        cand = np.array([[i / p, j / p] for i in range(p + 1) for j in range(p + 1 - i)])
    vals = np.stack([O.shape_functions(SHAPE[ct], p, x)[0] for x in cand])  # (cand, nl)
    pos = np.zeros((vals.shape[1], 2))
    for a in range(vals.shape[1]):
        k = np.flatnonzero(np.abs(vals[:, a] - 1.0) < 1e-12)
        assert k.size == 1
        assert np.abs(np.delete(vals[k[0]], a)).max() < 1e-12
        pos[a] = cand[k[0]]
    _NODES[(ct, p)] = pos
    return pos


def _edge_param(ct, p, e):
    """(slot, t) of the nodes on local edge e, t = position from its first vertex."""
    va, vb = VREF[ct][EDGES[ct][e][0]], VREF[ct][EDGES[ct][e][1]]
    out = []
    for a, x in enumerate(_nodes(ct, p)):
        d = x - va
        L = vb - va
        t = d @ L / (L @ L)
        if abs(d[0] * L[1] - d[1] * L[0]) < 1e-12 and -1e-12 <= t <= 1 + 1e-12:
            out.append((a, t))
    return out


def _constraint(ct, p, eord):
    """C[a, m]: value at node a of slot function m."""
    nl = len(_nodes(ct, p))
    C = np.eye(nl)
    for e, q in enumerate(eord[:len(EDGES[ct])]):
        if q == 0:
            continue
        on = sorted(_edge_param(ct, p, e), key=lambda at: at[1])  # vertex, interior..., vertex
        inner = [a for a, t in on[1:-1]]
        masters = [on[0][0]] + inner[:q - 1] + [on[-1][0]]         # the edge's dofs j = 0..q
        nodes_q = np.arange(q + 1) / q
        for a, t in on[1:-1]:
            C[a] = 0.0
            for j, mslot in enumerate(masters):
                y = np.zeros(q + 1)
                y[j] = 1.0
                C[a, mslot] = lagrange(nodes_q, y)(t)
    return C


def _cells(m):
    return [m.cell_nodes(c) for c in range(m.ncells)]


@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_one_order_is_the_reference_space(p):
    m, _ = _mesh(6, seed=p)
    om = _om(m)
    order = np.full(m.ncells, p, dtype=np.int32)
    dofs, eord, n = O.order_space(om, order)
    ig, off, n2 = O.dofhandler(om, p)
    assert n == n2 and not eord.any()
    for c in range(m.ncells):
        nl = (p + 1) * (p + 2) // 2 if m.celltypes[c] == TRI else (p + 1) ** 2
        assert np.array_equal(dofs[c, :nl], ig[off[c]:off[c] + nl] - 1)
        assert np.all(dofs[c, nl:] == -1)
    eta = np.random.default_rng(p).uniform(0.5, 2.0, m.ncells)
    deg = max(2 * p, 4)
    a = O.assemble_bilinear_orders(om, order, ["stiffness", "mass"], eta, degree=deg)
    b = O.assemble_bilinear(om, p, ["stiffness", "mass"], eta=eta, degree=deg)
    for x, y in zip(a[:3], b[:3]):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_traces_agree_on_every_interior_edge(seed):
    m, blocks = _mesh(8, seed=seed)
    order = _orders(blocks, seed)
    dofs, eord, n = O.order_space(_om(m), order)
    assert eord.any()
    used = np.unique(dofs[dofs >= 0])
    assert np.array_equal(used, np.arange(n))  # every dof id is used, none skipped
    cells = _cells(m)
    owners = {}
    for c, nodes in enumerate(cells):
        ct = int(m.celltypes[c])
        for e, (i, j) in enumerate(EDGES[ct]):
            owners.setdefault(tuple(sorted((nodes[i], nodes[j]))), []).append((c, e))
    ts = np.array([0.13, 0.37, 0.5, 0.71, 0.94])
    checked = 0
    for key, side in owners.items():
        if len(side) != 2:
            continue
        traces = []
        for c, e in side:
            ct, p = int(m.celltypes[c]), int(order[c])
            C = _constraint(ct, p, eord[c])
            ia, ib = EDGES[ct][e]
            t = ts if cells[c][ia] == key[0] else 1.0 - ts  # parameter from the edge's lower node id
            va, vb = VREF[ct][ia], VREF[ct][ib]
            phi = np.stack([O.shape_functions(SHAPE[ct], p, va + s * (vb - va))[0] for s in t])
            tr = phi @ C  # (points, slots): trace of each slot function
            d = {}
            for slot, g in enumerate(dofs[c, :C.shape[0]]):
                if g >= 0 and np.abs(tr[:, slot]).max() > 1e-13:
                    d[int(g)] = tr[:, slot]
            traces.append(d)
        assert traces[0].keys() == traces[1].keys(), (key, side)
        for g in traces[0]:
            assert np.abs(traces[0][g] - traces[1][g]).max() < 1e-12
        checked += 1
    assert checked > 100


@pytest.mark.parametrize("seed", [3, 4])
def test_embedding_in_the_uniform_highest_order_space(seed):
    m, blocks = _mesh(6, seed=seed)
    om = _om(m)
    order = _orders(blocks, seed, probs=(0.25, 0.25, 0.25, 0.25))
    pmax = int(order.max())
    deg = 2 * pmax
    eta = np.random.default_rng(seed).uniform(0.5, 2.0, m.ncells)
    dofs, eord, n = O.order_space(om, order)
    ig, off, nmax = O.dofhandler(om, pmax)
    P = np.full((nmax, n), np.nan)
    for c in range(m.ncells):
        ct, p = int(m.celltypes[c]), int(order[c])
        C = _constraint(ct, p, eord[c])
        X = _nodes(ct, pmax)
        I = np.stack([O.shape_functions(SHAPE[ct], p, x)[0] for x in X])  # (pmax nodes, p dofs)
        loc = I @ C
        gmax = ig[off[c]:off[c] + X.shape[0]] - 1
        for s, g in enumerate(dofs[c, :C.shape[0]]):
            if g < 0:
                continue
            prev = P[gmax, g]
            col = loc[:, s]
            ok = np.isnan(prev) | (np.abs(prev - col) < 1e-12)
            assert ok.all(), "the space's functions must agree at shared nodes"
            P[gmax, g] = col
    P = np.nan_to_num(P)  # nodes of cells where the dof's function is absent: zero there
    # every uniform node lies in some cell, so each column was set where its function lives
    terms = ["stiffness", "mass"]
    cp, rv, nz, nn = O.assemble_bilinear(om, pmax, terms, eta=eta, degree=deg)
    Amax = O.csc_to_scipy(cp, rv, nz, nn).toarray()
    cp2, rv2, nz2, n2 = O.assemble_bilinear_orders(om, order, terms, eta, degree=deg)
    A = O.csc_to_scipy(cp2, rv2, nz2, n2).toarray()
    ref = P.T @ Amax @ P
    err = np.abs(A - ref).max() / np.abs(ref).max()
    print(f"orders {np.bincount(order)[1:]}, pmax {pmax}: |A - PᵀAP| / max = {err:.2e}")
    assert err < 1e-12
    # every nonzero of PᵀAP is in the assembled pattern
    assert not np.any((np.abs(ref) > 1e-12 * np.abs(ref).max()) & (A == 0))


def test_constants_and_area():
    m, blocks = _mesh(10, seed=7)
    om = _om(m)
    order = _orders(blocks, 7)
    cpK, rvK, K, n = O.assemble_bilinear_orders(om, order, ["stiffness"])
    cpM, rvM, M, _ = O.assemble_bilinear_orders(om, order, ["mass"])
    Ks = O.csc_to_scipy(cpK, rvK, K, n)
    Ms = O.csc_to_scipy(cpM, rvM, M, n)
    one = np.ones(n)
    assert np.abs(Ks @ one).max() < 1e-12 * np.abs(K).max()
    assert abs(one @ (Ms @ one) - 1.0) < 1e-13
    b = O.assemble_linear_orders(om, order, "one")
    assert abs(b.sum() - 1.0) < 1e-13
    assert np.allclose(b, Ms @ one, rtol=0, atol=1e-15)
