"""Pins the CPU oracle against the reference's own known answers.

Every check here comes from Bcube.jl's test suite (integer fixtures copied as
data into tests/golden/) or from mathematics that fixes the answer (closed-form
integrals, Gauss exactness, an analytic PDE solution), chosen so that a dropped
term, a wrong sign or a transposed index fails at least one of them.
"""
import json
import os

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle import oracle as O

HERE = os.path.dirname(__file__)
CASES = json.load(open(os.path.join(HERE, "golden", "dofhandler_cases.json")))


def _mesh_from_case(m):
    if m["kind"] == "rectangle":
        return O.rectangle_mesh(m["nx"], m["ny"], order=m["order"])
    nodes = np.array(m["nodes"], float)
    cells = m["cells"]
    ptr = np.cumsum([0] + [len(c) for c in cells])
    c2n = np.concatenate([np.array(c) - 1 for c in cells])
    return O.OMesh(m["dim"], nodes, [O.CT[t] for t in m["celltypes"]], ptr, c2n)


@pytest.mark.parametrize("case", CASES["cases"], ids=[c["name"] for c in CASES["cases"]])
def test_dofhandler_fixtures(case):
    """test/fespace/test_dofhandler.jl integer vectors, exact."""
    mesh = _mesh_from_case(case["mesh"])
    iglob, offset, _ = O.dofhandler(mesh, case["order"], case["continuous"])
    ends = list(offset[1:]) + [iglob.size]
    for k, want in case["expect"].items():
        c = int(k) - 1
        assert iglob[offset[c]:ends[c]].tolist() == want


@pytest.mark.parametrize("t", CASES["densify"])
def test_densify_fixtures(t):
    assert O.densify(t["in"]).tolist() == t["out"]


def test_rectangle_numbering():
    """mesh_generator.jl:253-332: node (iy-1)*nx+ix, cell c2n CCW from (ix,iy)."""
    m = O.rectangle_mesh(4, 3)
    c2n = m.c2n.reshape(-1, 4)
    assert c2n[0].tolist() == [0, 1, 5, 4]
    assert c2n[4].tolist() == [5, 6, 10, 9]  # cell (ix=2, iy=2)
    xy = m.coords.reshape(-1, 2)
    assert np.allclose(xy[5], [1 / 3, 0.5])


def test_hexa_numbering():
    m = O.hexa_mesh(3, 3, 3)
    c2n = m.c2n.reshape(-1, 8)
    assert c2n[0].tolist() == [0, 1, 4, 3, 9, 10, 13, 12]
    assert m.coords.reshape(-1, 3)[13].tolist() == [0.5, 0.5, 0.5]


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6])
def test_gauss_legendre_vs_numpy(n):
    w, x = O.quadrature("Line", 2 * n - 1)
    xr, wr = np.polynomial.legendre.leggauss(n)
    assert len(w) == n
    np.testing.assert_allclose(x[:, 0], xr, rtol=0, atol=2e-16)
    np.testing.assert_allclose(w, wr, rtol=1e-14, atol=0)  # numpy itself is a few ulp off at n=6


def test_quadrature_sizes_and_sums():
    """test/quadrature/test_quadrature.jl:23-56."""
    w, _ = O.quadrature("Line", 2)
    assert len(w) == 2 and np.isclose(w.sum(), 2)
    w, _ = O.quadrature("Line", 1)
    assert len(w) == 1 and np.isclose(w.sum(), 2)
    w, _ = O.quadrature("Line", 2, qtype=1)
    assert len(w) == 3 and np.isclose(w.sum(), 1)
    w, _ = O.quadrature("Line", 3)
    assert len(w) == 2
    w, x = O.quadrature("Square", 2)
    assert len(w) == 4 and np.isclose(w.sum(), 4) and x.shape == (4, 2)
    w, x = O.quadrature("Cube", 2)
    assert len(w) == 8 and np.isclose(w.sum(), 8) and x.shape == (8, 3)


def test_tensor_rule_order_first_index_fastest():
    """quadrature.jl:455-487: x = (x1[i], x1[j]) with i fastest."""
    w, x = O.quadrature("Square", 3)
    g = 1 / np.sqrt(3)
    assert np.allclose(x, [[-g, -g], [g, -g], [-g, g], [g, g]])
    w, x = O.quadrature("Cube", 5)
    assert len(w) == 27 and np.allclose(x[1], [0, -np.sqrt(0.6), -np.sqrt(0.6)])


@pytest.mark.parametrize("deg", [1, 2, 3, 4, 5, 7])
def test_cube_quadrature_exactness(deg):
    w, x = O.quadrature("Cube", deg)
    rng = np.random.default_rng(deg)
    for _ in range(5):
        p = rng.integers(0, deg + 1, size=3)
        exact = np.prod([(1 - (-1) ** (k + 1)) / (k + 1) for k in p])
        got = np.sum(w * np.prod(x ** p, axis=1))
        assert np.isclose(got, exact, atol=1e-14)


def test_square_q1_values_and_gradients():
    """test/interpolation/test_lagrange.jl:264-278 (dof order is tensor, not CCW)."""
    v, _ = O.shape_functions("Square", 1, [-1, -1]); assert np.isclose(v[0], 1)
    v, _ = O.shape_functions("Square", 1, [1, -1]); assert np.isclose(v[1], 1)
    v, _ = O.shape_functions("Square", 1, [1, 1]); assert np.isclose(v[3], 1)
    v, _ = O.shape_functions("Square", 1, [-1, 1]); assert np.isclose(v[2], 1)
    v, _ = O.shape_functions("Square", 1, [np.pi, np.e]); assert np.isclose(v.sum(), 1)
    v, g = O.shape_functions("Square", 1, [0, 0])
    assert v.tolist() == [0.25] * 4
    assert np.allclose(g, np.array([[-1, -1], [1, -1], [-1, 1], [1, 1]]) / 4)


def test_square_q2_values():
    """test_lagrange.jl:290-310."""
    v, _ = O.shape_functions("Square", 2, [-1, -1]); assert np.isclose(v[0], 1)
    v, _ = O.shape_functions("Square", 2, [1, -1]); assert np.isclose(v[2], 1)
    v, _ = O.shape_functions("Square", 2, [1, 1]); assert np.isclose(v[8], 1)
    v, _ = O.shape_functions("Square", 2, [-1, 1]); assert np.isclose(v[6], 1)
    v, _ = O.shape_functions("Square", 2, [0, 0])
    assert v.tolist() == [0, 0, 0, 0, 1, 0, 0, 0, 0]


@pytest.mark.parametrize("shape,order", [("Square", 1), ("Square", 2), ("Square", 3),
                                         ("Cube", 1), ("Cube", 2), ("Cube", 3)])
def test_nodal_delta_and_gradient_fd(shape, order):
    """test_lagrange.jl:8-45 nodal property + finite-difference gradient check."""
    d = 2 if shape == "Square" else 3
    n = order + 1
    nodes1 = np.linspace(-1, 1, n)
    grids = np.meshgrid(*([nodes1] * d), indexing="ij")
    pts = np.stack([g.reshape(-1, order="F") for g in grids], axis=1)  # first index fastest
    for k, p in enumerate(pts):
        v, _ = O.shape_functions(shape, order, p)
        e = np.zeros(len(v)); e[k] = 1
        np.testing.assert_allclose(v, e, atol=1e-13)
    rng = np.random.default_rng(0)
    xi = rng.uniform(-1, 1, d)
    _, g = O.shape_functions(shape, order, xi)
    h = 1e-6
    for a in range(d):
        dp = xi.copy(); dp[a] += h
        dm = xi.copy(); dm[a] -= h
        fd = (O.shape_functions(shape, order, dp)[0] - O.shape_functions(shape, order, dm)[0]) / (2 * h)
        np.testing.assert_allclose(g[:, a], fd, atol=1e-8)


def test_hexa8_jacobian_scaled_cube():
    """test_integration.jl:476-503: J = s*I on a scaled/translated cube."""
    base = O.hexa_mesh(2, 2, 2, -1, 1, -1, 1, -1, 1)
    X = base.coords.reshape(-1, 3)[base.c2n]
    rng = np.random.default_rng(1)
    for _ in range(5):
        s = 10 * rng.random()
        t = rng.random(3)
        J, Ji, det, _ = O.geometry("Hexa8", (X + t) * s, [0.0, 0.0, 0.0])
        np.testing.assert_allclose(J, s * np.eye(3), rtol=1e-14)
        np.testing.assert_allclose(det * 8, 8 * s ** 3, rtol=1e-14)
        np.testing.assert_allclose(Ji @ J, np.eye(3), atol=1e-14)


def test_quad_mass_stiffness_closed_forms():
    """test_integration.jl:224-256 on one [0,1]^2 quad, Quadrature(2)."""
    mesh = O.rectangle_mesh(2, 2)
    colptr, rowval, nzval, n = O.assemble_bilinear(mesh, 1, ["mass"], degree=2)
    M = O.csc_to_scipy(colptr, rowval, nzval, n).toarray()
    # dof 1 <-> vertex 1, dof 2 <-> vertex 2 (mass = ΔxΔy/9 on the diagonal, /18 adjacent)
    assert np.isclose(M[0, 0], 1 / 9) and np.isclose(M[0, 1], 1 / 18) and np.isclose(M[1, 1], 1 / 9)
    colptr, rowval, nzval, n = O.assemble_bilinear(mesh, 1, ["stiffness"], degree=2)
    K = O.csc_to_scipy(colptr, rowval, nzval, n).toarray()
    assert np.isclose(K[0, 0], (1 + 1) / 3)


def test_csc_keeps_structure_and_sums_duplicates():
    mesh = O.rectangle_mesh(3, 3)
    colptr, rowval, nzval, n = O.assemble_bilinear(mesh, 1, ["stiffness", "mass"], degree=2)
    assert n == 9 and colptr[-1] - 1 == 49  # 3x3 node grid, 9-point stencil pattern
    for j in range(n):  # rows sorted within each column
        r = rowval[colptr[j] - 1:colptr[j + 1] - 1]
        assert np.all(np.diff(r) > 0)
    A = O.csc_to_scipy(colptr, rowval, nzval, n)
    K = O.csc_to_scipy(*O.assemble_bilinear(mesh, 1, ["stiffness"], degree=2))
    M = O.csc_to_scipy(*O.assemble_bilinear(mesh, 1, ["mass"], degree=2))
    assert abs(A - K - M).max() < 1e-15
    assert abs(K @ np.ones(n)).max() < 1e-14  # constants are in the kernel
    assert np.isclose(M.sum(), 1.0)  # Σ mass = |Ω|


def _dof_coords(mesh, order, iglob, offset, ndofs):
    """Physical location of every Lagrange dof (Hexa8/Quad4 mapping of the nodes)."""
    d = mesh.dim
    n1 = np.linspace(-1, 1, order + 1)
    grids = np.meshgrid(*([n1] * d), indexing="ij")
    ref = np.stack([g.reshape(-1, order="F") for g in grids], axis=1)
    X = mesh.coords.reshape(-1, d)
    k = 8 if d == 3 else 4
    ct = "Hexa8" if d == 3 else "Quad4"
    out = np.zeros((ndofs, d))
    for c in range(mesh.ncells):
        cn = X[mesh.c2n[c * k:(c + 1) * k]]
        for ld, xi in enumerate(ref):
            out[iglob[offset[c] + ld] - 1] = O.geometry(ct, cn, xi)[3]
    return out


def test_heat_solver_3d_q2_analytic():
    """test_assembler.jl:457-509: hexa_mesh(5,5,5) on [-1/2,1/2]^3, Q2, λ=100,
    q=1500, T=260 on xmin; the quadratic analytic solution is reproduced to 2e-14."""
    mesh = O.hexa_mesh(5, 5, 5, -0.5, 0.5, -0.5, 0.5, -0.5, 0.5)
    lam, q, order = 100.0, 1500.0, 2
    eta = np.full(mesh.ncells, lam)
    A = O.csc_to_scipy(*O.assemble_bilinear(mesh, order, ["stiffness"], eta=eta, degree=5)).tocsr()
    b = O.assemble_linear(mesh, order, "const", q, degree=5)
    iglob, offset, n = O.dofhandler(mesh, order, True)
    xd = _dof_coords(mesh, order, iglob, offset, n)
    T = lambda x: 260.0 + (q / lam) * (x + 0.5) * (1.0 - 0.5 * (x + 0.5))
    bnd = np.isclose(xd[:, 0], -0.5)
    u = np.zeros(n)
    u[bnd] = 260.0
    free = ~bnd
    rhs = b[free] - A[free][:, bnd] @ u[bnd]
    u[free] = spla.spsolve(A[free][:, free].tocsc(), rhs)
    Tref = T(xd[:, 0])
    err = np.abs(u - Tref).max() / np.abs(Tref).max()
    assert err < 2e-14


def test_linear_source_exact_p1_quad():
    """∫ v dΩ on a unit-square Q1 mesh: each vertex gets its patch area / 4."""
    mesh = O.rectangle_mesh(3, 3)
    b = O.assemble_linear(mesh, 1, "one", degree=2)
    np.testing.assert_allclose(sorted(b), sorted([1 / 16] * 4 + [1 / 8] * 4 + [1 / 4]), atol=1e-16)


def test_threaded_oracle_equals_serial():
    """The all-core CPU baseline (OpenMP COO build and sparse(), parallel
    element vectors of the linear form, parallel DG faces and cells) returns
    the serial oracle's results bit for bit, for any thread count."""
    import bcubehip as bh
    host = bh.BcubeBackendHostOnly()
    n = 5
    o = O.hexa_mesh(n + 1, n + 1, n + 1)
    X = o.coords.reshape(-1, 3)
    inner = np.all((X > 1e-12) & (X < 1 - 1e-12), axis=1)
    X[inner] += np.random.default_rng(0).uniform(-0.15, 0.15, X.shape)[inner] / n
    eta = np.random.default_rng(1).uniform(0.5, 2.0, o.ncells)
    ref = O.assemble_bilinear(o, 2, ["stiffness", "mass"], eta=eta, degree=5)
    m, _ = bh.mixed_tri_quad_mesh(12, block=4, p_split=0.5, seed=3, backend=host)
    om = O.OMesh(2, m.coords, m.celltypes, m.c2n_ptr, m.c2n)
    ref_mixed = O.assemble_bilinear(om, 3, ["stiffness", "mass"], degree=6)
    o2 = O.rectangle_mesh(17, 17)
    bref = O.assemble_linear(o2, 1, "sinsin", degree=2)
    h27 = bh.hexa27_mesh(4, 4, 4, backend=host)
    bh.sinusoidal_map(h27, 0.05, seed=0)
    o27 = O.OMesh(h27.dim, h27.coords, h27.celltypes, h27.c2n_ptr, h27.c2n)
    u = np.random.default_rng(2).normal(size=(27 * 27, 3))
    rref = O.dg_transport(o27, 2, [1.0, 0.5, 0.25], u, 5)
    for threads in (1, 3, 4):
        got, t = O.assemble_bilinear_timed(o, 2, ["stiffness", "mass"], eta=eta, degree=5,
                                           threads=threads, reps=1)
        assert t.shape == (1, 2) and (t > 0).all()
        for a, b in zip(got, ref):
            assert np.array_equal(a, b)
        got, _ = O.assemble_bilinear_timed(om, 3, ["stiffness", "mass"], degree=6,
                                           threads=threads, reps=1)
        for a, b in zip(got, ref_mixed):
            assert np.array_equal(a, b)
        b, _ = O.assemble_linear_timed(o2, 1, "sinsin", degree=2, threads=threads, reps=1)
        assert np.array_equal(b, bref)
        r, _ = O.dg_transport_timed(o27, 2, [1.0, 0.5, 0.25], u, 5, threads=threads, reps=1)
        assert np.array_equal(r, rref)


def _bubble_map(X, amp=0.04):
    """x_d += amp sin(πx) sin(πy) (sin(πz)) (1 + d): a smooth map that moves no
    node of the unit box's boundary, so the domain stays the box."""
    Y = X.copy()
    b = np.prod(np.sin(np.pi * X), axis=1)
    for d in range(X.shape[1]):
        Y[:, d] += amp * (1 + d) * b
    return Y


def _interior_patch_error(A, xd, u_lin):
    """K u = 0 in the interior with u = u_lin on the boundary dofs: the
    interior solution's max error against u_lin at the interior dofs."""
    inner = np.all((xd > 1e-12) & (xd < 1 - 1e-12), axis=1)
    K = A.tocsr()
    KII = K[inner][:, inner].tocsc()
    KIB = K[inner][:, ~inner]
    u = u_lin(xd)
    ui = spla.spsolve(KII, -KIB @ u[~inner])
    return np.abs(ui - u[inner]).max() / np.abs(u).max()


def _q2_dof_points(om, order, ndofs, iglob_offset):
    """Physical points of the Q2 dofs of a curved mesh: the dof at reference
    point ξ of a cell sits at the geometric node with that ξ (the Q2 dofs and
    the Quad9/Hexa27 nodes share their reference points)."""
    iglob, offset = iglob_offset
    d = om.dim
    X = om.coords.reshape(-1, d)
    k = om.c2n_ptr[1] - om.c2n_ptr[0]
    if d == 2:
        ref = np.array([[-1, -1], [1, -1], [1, 1], [-1, 1], [0, -1], [1, 0], [0, 1], [-1, 0],
                        [0, 0]])
    else:
        import bcubehip.mesh as bm
        ref = bm._H27_REF
    lex = [tuple(r) for r in (np.stack(np.meshgrid(*([[-1, 0, 1]] * d), indexing="ij"), -1)
                              .reshape(-1, d)[:, ::-1])]
    node_of = [int(np.flatnonzero((ref == np.array(p)).all(1))[0]) for p in lex]
    pts = np.zeros((ndofs, d))
    for c in range(om.ncells):
        nodes = om.c2n[c * k:(c + 1) * k]
        for loc in range(len(lex)):
            pts[iglob[offset[c] + loc] - 1] = X[nodes[node_of[loc]]]
    return pts


@pytest.mark.parametrize("dim", [2, 3])
def test_curved_q2_patch_test(dim):
    """The isoparametric Q2 stiffness on curved Quad9 / Hexa27 cells (a bubble
    map of the unit box) reproduces a linear field: with linear Dirichlet data
    the interior solution is that field, to rounding.  A wrong Jacobian, a
    transposed J⁻¹ or a wrong node order in the quadratic map fails it."""
    import bcubehip as bh
    host = bh.BcubeBackendHostOnly()
    n = 4
    if dim == 2:
        m = bh.rectangle_mesh(n + 1, n + 1, order=2, backend=host)
    else:
        m = bh.hexa27_mesh(n + 1, n + 1, n + 1, backend=host)
    m.set_nodes(_bubble_map(m.nodes()))
    om = O.OMesh(m.dim, m.coords, m.celltypes, m.c2n_ptr, m.c2n)
    cp, rv, nz, nd = O.assemble_bilinear(om, 2, ["stiffness"], degree=5 if dim == 3 else 6)
    A = O.csc_to_scipy(cp, rv, nz, nd)
    iglob, offset, n_d = O.dofhandler(om, 2)
    assert n_d == nd
    xd = _q2_dof_points(om, 2, nd, (iglob, offset))
    u_lin = (lambda x: 1.0 + 2.0 * x[:, 0] - 3.0 * x[:, 1]) if dim == 2 else \
        (lambda x: 1.0 + 2.0 * x[:, 0] - 3.0 * x[:, 1] + 0.5 * x[:, 2])
    err = _interior_patch_error(A, xd, u_lin)
    assert err < 1e-11, err
    # the cells are really curved: first-edge midpoint nodes (CGNS node 4 of a
    # Quad9, 8 of a Hexa27) lie off the chord of their two vertices
    Xn = om.coords.reshape(-1, dim)
    k = om.c2n_ptr[1] - om.c2n_ptr[0]
    c2n = om.c2n.reshape(-1, k)
    mid = c2n[:, 4 if dim == 2 else 8]
    off = np.linalg.norm(Xn[mid] - 0.5 * (Xn[c2n[:, 0]] + Xn[c2n[:, 1]]), axis=1)
    assert off.max() > 1e-3


def test_jittered_hexa8_q2_patch_test():
    """The jittered Hexa8 Q2 stiffness of the C2 workload reproduces a linear
    field (the trilinear map is affine on no cell here): pins the oracle's
    J⁻¹ orientation on non-affine cells directly."""
    n = 4
    o = O.hexa_mesh(n + 1, n + 1, n + 1)
    X = o.coords.reshape(-1, 3)
    inner = np.all((X > 1e-12) & (X < 1 - 1e-12), axis=1)
    X[inner] += np.random.default_rng(0).uniform(-0.15, 0.15, X.shape)[inner] / n
    cp, rv, nz, nd = O.assemble_bilinear(o, 2, ["stiffness"], degree=5)
    A = O.csc_to_scipy(cp, rv, nz, nd)
    # Q2 dof points on trilinear cells: the trilinear map at the lattice points
    nl = 27
    iglob, offset, _ = O.dofhandler(o, 2)
    v8 = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1], [-1, -1, 1], [1, -1, 1],
                   [1, 1, 1], [-1, 1, 1]], dtype=np.float64)
    lex = np.stack(np.meshgrid([-1.0, 0.0, 1.0], [-1.0, 0.0, 1.0], [-1.0, 0.0, 1.0],
                               indexing="ij"), -1).reshape(-1, 3)[:, ::-1]
    w = np.prod((1.0 + lex[:, None, :] * v8[None, :, :]) / 2.0, axis=2)
    xd = np.zeros((nd, 3))
    for c in range(o.ncells):
        xd[iglob[offset[c]:offset[c] + nl] - 1] = w @ X[o.c2n[8 * c:8 * c + 8]]
    err = _interior_patch_error(A, xd, lambda x: 0.5 - x[:, 0] + 2.0 * x[:, 1] + 3.0 * x[:, 2])
    assert err < 1e-11, err


def test_quad9_on_straight_cells_equals_quad4():
    """Quad9 cells whose edge and centre nodes sit at the bilinear map's points
    describe the Quad4 cells: same matrices (pins the quadratic map's node
    order against the bilinear one), on a jittered mesh."""
    import bcubehip as bh
    host = bh.BcubeBackendHostOnly()
    m4 = bh.rectangle_mesh(6, 5, backend=host)
    bh.jitter_nodes(m4, 0.2, 0.2, seed=3)
    X = m4.nodes()
    c4 = m4.c2n.reshape(-1, 4)
    mids = {}
    nodes = [tuple(p) for p in X]
    c9 = []
    for cell in c4:
        P = X[cell]
        row = list(cell)
        for a, b in ((0, 1), (1, 2), (2, 3), (3, 0)):
            key = (min(cell[a], cell[b]), max(cell[a], cell[b]))
            if key not in mids:
                mids[key] = len(nodes)
                nodes.append(tuple((P[a] + P[b]) / 2))
            row.append(mids[key])
        row.append(len(nodes))
        nodes.append(tuple(P.mean(0)))
        c9.append(row)
    o4 = O.OMesh(2, X, np.zeros(len(c4), np.int32), np.arange(len(c4) + 1) * 4, c4.reshape(-1))
    o9 = O.OMesh(2, np.array(nodes), np.full(len(c9), 1, np.int32),
                 np.arange(len(c9) + 1) * 9, np.array(c9).reshape(-1))
    for order in (1, 2):
        a = O.assemble_bilinear(o4, order, ["stiffness", "mass"], degree=4)
        b = O.assemble_bilinear(o9, order, ["stiffness", "mass"], degree=4)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
        assert np.abs(a[2] - b[2]).max() <= 1e-14 * np.abs(a[2]).max()
