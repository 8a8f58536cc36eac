"""Vector spaces (size > 1) and MultiFESpace on the host: the numbering of
the product (vector ids derived from the scalar numbering, MultiFESpace
mappings vectorised) against the reference's own vectors
(test/fespace/test_dofhandler.jl, as tests/golden/multifield_cases.json) and
against the oracle's restatement of the reference loops (DofHandler with
ncomponents, _build_mapping_AoS/SoA); the global CSC pattern of block
assembly bit-exact with the oracle's (assembler.jl:327-420).  No GPU needed."""
import json
import os

import numpy as np
import pytest

import bcubehip as B
from oracle import oracle as O

HOST = B.BcubeBackendHostOnly()
GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "multifield_cases.json")))
LAG = lambda p: B.FunctionSpace("Lagrange", p)  # noqa: E731


def _meshes(m):
    nodes = np.array(m["nodes"], float)
    cells = [np.array(c) - 1 for c in m["cells"]]
    ptr = np.cumsum([0] + [len(c) for c in cells])
    om = O.OMesh(m["dim"], nodes, [O.CT[t] for t in m["celltypes"]], ptr, np.concatenate(cells))
    bm = B.Mesh(nodes, m["celltypes"], [c.tolist() for c in cells], HOST)
    return om, bm


def _om(m):
    return O.OMesh(m.dim, m.coords, m.celltypes, m.c2n_ptr, m.c2n)


def _oracle_spaces(om, specs):
    nc = len(om.celltypes)
    out = []
    for sp in specs:
        ig, off, n = O.dofhandler_n(om, sp["order"], sp["size"], sp["continuous"])
        nd = np.diff(np.append(off[:, 0], ig.size)) // sp["size"]
        out.append((ig, off, n, nd))
    return out, nc


@pytest.mark.parametrize("case", GOLD["cases"], ids=[c["name"] for c in GOLD["cases"]])
def test_reference_vectors(case):
    om, bm = _meshes(case["mesh"])
    specs = case["spaces"]
    spaces = [B.TrialFESpace(LAG(s["order"]), bm, "continuous" if s["continuous"] else
                             "discontinuous", size=s["size"]) for s in specs]
    ospaces, nc = _oracle_spaces(om, specs)
    if "component_dofs" in case:
        (ig, off, n, nd), = ospaces
        for cell, comps in case["component_dofs"].items():
            c = int(cell) - 1
            for k, want in enumerate(comps):
                assert spaces[0].dhl.get_dof(c, k).tolist() == want
                assert ig[off[c, k]:off[c, k] + nd[c]].tolist() == want
    if "mapped" in case:
        M = B.MultiFESpace(*spaces, arrayOfStruct=case["aos"])
        dofs_of = [lambda c, s=s, i=i: O.cell_dofs(s[0], s[1], specs[i]["size"], s[3], c)
                   for i, s in enumerate(ospaces)]
        _, omap = O.multi_mapping(dofs_of, [s[2] for s in ospaces], nc, case["aos"])
        for si, per_cell in case["mapped"].items():
            i = int(si)
            for cell, want in per_cell.items():
                c = int(cell) - 1
                assert M.get_mapping(i)[spaces[i].get_dofs(c) - 1].tolist() == want
                assert omap[i][dofs_of[i](c) - 1].tolist() == want


def _mesh(kind):
    if kind == "quad":
        m = B.rectangle_mesh(5, 4, backend=HOST)
    elif kind == "mixed":
        m, _ = B.mixed_tri_quad_mesh(6, block=2, p_split=0.5, seed=3, backend=HOST)
    elif kind == "tri":
        m, _ = B.mixed_tri_quad_mesh(4, block=4, p_split=1.0, seed=1, backend=HOST)
    else:
        m = B.hexa_mesh(3, 3, 2, backend=HOST)
    return m


@pytest.mark.parametrize("kind,p,ncomp", [("quad", 1, 2), ("quad", 3, 3), ("mixed", 2, 2),
                                          ("tri", 3, 2), ("hexa", 1, 3), ("hexa", 2, 3)])
@pytest.mark.parametrize("continuous", [True, False])
def test_vector_numbering_matches_oracle(kind, p, ncomp, continuous):
    m = _mesh(kind)
    U = B.TrialFESpace(LAG(p), m, "continuous" if continuous else "discontinuous", size=ncomp)
    ig, off, n = O.dofhandler_n(_om(m), p, ncomp, continuous)
    assert U.get_ndofs() == n
    assert np.array_equal(U.dhl.iglob, ig)
    assert np.array_equal(U.dhl.offset, off[:, 0])
    # the component map is a bijection onto 0..n-1
    cm = U.component_map
    assert np.array_equal(np.sort(cm.reshape(-1)), np.arange(n))


@pytest.mark.parametrize("aos", [True, False])
def test_multi_mapping_matches_oracle_loops(aos):
    m = _mesh("mixed")
    specs = [{"order": 2, "size": 1, "continuous": True}, {"order": 2, "size": 2, "continuous": True},
             {"order": 1, "size": 1, "continuous": False}]
    spaces = [B.TrialFESpace(LAG(s["order"]), m, "continuous" if s["continuous"] else
                             "discontinuous", size=s["size"]) for s in specs]
    M = B.MultiFESpace(*spaces, arrayOfStruct=aos)
    ospaces, nc = _oracle_spaces(_om(m), specs)
    dofs_of = [lambda c, s=s, i=i: O.cell_dofs(s[0], s[1], specs[i]["size"], s[3], c)
               for i, s in enumerate(ospaces)]
    n, omap = O.multi_mapping(dofs_of, [s[2] for s in ospaces], nc, aos)
    assert M.get_ndofs() == n
    for i in range(3):
        assert np.array_equal(M.get_mapping(i), omap[i])


def _vector_case(m, p, size):
    U = B.TrialFESpace(LAG(p), m, size=size)
    dΩ = B.Measure(B.CellDomain(m), 2 * p)
    a = lambda u, v: (B.integrate(B.ddot(B.grad(u), B.grad(v)), dΩ)  # noqa: E731
                      + B.integrate(B.dot(u, v), dΩ))
    return U, B.TestFESpace(U), a, [size], {(0, 0): (["stiffness", "mass"], None)}


def _multi_case(m, p, aos):
    sp = [B.TrialFESpace(LAG(p), m), B.TrialFESpace(LAG(p), m, size=2), B.TrialFESpace(LAG(p), m)]
    U = B.MultiFESpace(*sp, arrayOfStruct=aos)
    V = B.MultiFESpace(*[B.TestFESpace(s) for s in sp], arrayOfStruct=aos)
    dΩ = B.Measure(B.CellDomain(m), 2 * p)

    def a(u, v):
        return (B.integrate(u[0] * v[0], dΩ) + B.integrate(B.ddot(B.grad(u[1]), B.grad(v[1])), dΩ)
                + B.integrate(u[0] * v[2] + 2.0 * B.dot(B.grad(u[2]), B.grad(v[2])), dΩ))
    blocks = {(0, 0): (["mass"], None), (1, 1): (["stiffness"], None), (2, 0): (["mass"], None),
              (2, 2): (["stiffness"], np.full(m.ncells, 2.0))}
    return U, V, a, [1, 2, 1], blocks


@pytest.mark.parametrize("kind,p", [("quad", 2), ("mixed", 1), ("hexa", 1)])
def test_vector_pattern_bit_exact(kind, p):
    m = _mesh(kind)
    U, V, a, sizes, blocks = _vector_case(m, p, 3 if kind == "hexa" else 2)
    plan = B.MultiBilinearPlan(a, U, V)
    cp, rv = plan.csc_structure()
    ocp, orv, _, n = O.assemble_blocks(_om(m), p, sizes, blocks, degree=2 * p)
    assert plan.nrows == n == U.get_ndofs()
    assert np.array_equal(cp + 1, ocp) and np.array_equal(rv + 1, orv)


@pytest.mark.parametrize("aos", [True, False])
def test_multifield_pattern_bit_exact(aos):
    m = _mesh("mixed")
    U, V, a, sizes, blocks = _multi_case(m, 2, aos)
    plan = B.MultiBilinearPlan(a, U, V)
    cp, rv = plan.csc_structure()
    ocp, orv, _, n = O.assemble_blocks(_om(m), 2, sizes, blocks, aos=aos, degree=4)
    assert plan.nrows == n == U.get_ndofs()
    assert np.array_equal(cp + 1, ocp) and np.array_equal(rv + 1, orv)
    # a coupling listed one way only: (test 2, trial 0) present, (test 0, trial 2) absent
    A = plan.to_csc(np.ones(plan.nnz))
    m0, m2 = U.get_mapping(0) - 1, U.get_mapping(2) - 1
    assert A[m2[0], m0[0]] == 1.0 and A[m0[0], m2[0]] == 0.0


def test_unsupported_couplings_raise():
    m = _mesh("quad")
    sp = [B.TrialFESpace(LAG(1), m), B.TrialFESpace(LAG(1), m, size=2)]
    U = B.MultiFESpace(*sp)
    V = B.MultiFESpace(*[B.TestFESpace(s) for s in sp])
    dΩ = B.Measure(B.CellDomain(m), 2)
    with pytest.raises(NotImplementedError):  # u_xy⋅∇(v_p): a vector trial, a scalar test space
        B.MultiBilinearPlan(lambda u, v: B.integrate(B.dot(u[1], B.grad(v[0])), dΩ), U, V)
    W = B.MultiFESpace(B.TrialFESpace(LAG(1), m), B.TrialFESpace(LAG(2), m))
    with pytest.raises(NotImplementedError):  # two orders: rectangular blocks
        B.MultiBilinearPlan(lambda u, v: B.integrate(u[0] * v[0], dΩ), W, W)
    Uv = B.TrialFESpace(LAG(1), m, size=2)
    with pytest.raises(NotImplementedError):  # ∇u⋅v: a matrix against a vector
        B.MultiBilinearPlan(lambda u, v: B.integrate(B.dot(B.grad(u), v), dΩ), Uv,
                            B.TestFESpace(Uv))


def test_face_measures_on_vector_spaces_raise():
    """Vector and multi-field assembly integrates over cells only: a face
    measure must be refused, not assembled as a second cell integral."""
    m = _mesh("quad")
    Uv = B.TrialFESpace(LAG(1), m, size=2)
    Vv = B.TestFESpace(Uv)
    dΩ = B.Measure(B.CellDomain(m), 2)
    for dom in (B.BoundaryFaceDomain(m), B.InteriorFaceDomain(m)):
        dΓ = B.Measure(dom, 2)
        with pytest.raises(NotImplementedError):
            B.MultiBilinearPlan(lambda u, v: B.integrate(B.dot(u, v), dΩ)
                                + B.integrate(B.dot(u, v), dΓ), Uv, Vv)
        g = B.PhysicalFunction(lambda x: np.ones((x.shape[0], 2)))
        with pytest.raises(NotImplementedError):
            B.assemble_linear(lambda v: B.integrate(B.dot(g, v), dΓ), Vv)


def test_boundary_face_domain_refuses_labels():
    """BoundaryFaceDomain(mesh, label) selects a subset of faces in the
    reference; the face kernels take all of them, so a label raises."""
    m = _mesh("quad")
    with pytest.raises(NotImplementedError):
        B.BoundaryFaceDomain(m, "xmin")
    bc = B.PeriodicBCType(B.Translation([1.0, 0.0]), ("xmax",), ("xmin",))
    dΓp = B.Measure(B.BoundaryFaceDomain(m, bc), 2)
    U = B.TrialFESpace(LAG(1), m, "discontinuous")
    with pytest.raises(NotImplementedError):  # periodic faces in a face-term form
        B.assemble_bilinear(lambda u, v: B.integrate(B.side_n(u) * B.side_n(v), dΓp), U,
                            B.TestFESpace(U))


def test_dot_of_vector_gradients_is_the_frobenius_product():
    """bench_assemble.jl:10-28 assembles ∫(∇(u) ⋅ ∇(v))dΩ on a size-2 space:
    Julia's dot of two matrices contracts both indices, the form of ∇u ⊡ ∇v."""
    m = _mesh("quad")
    Uv = B.TrialFESpace(LAG(1), m, size=2)
    dΩ = B.Measure(B.CellDomain(m), 3)
    a = B.MultiBilinearPlan(lambda u, v: B.integrate(B.dot(B.grad(u), B.grad(v)), dΩ), Uv,
                            B.TestFESpace(Uv))
    b = B.MultiBilinearPlan(lambda u, v: B.integrate(B.ddot(B.grad(u), B.grad(v)), dΩ), Uv,
                            B.TestFESpace(Uv))
    assert [(t.form, t.coef_kind, t.scale) for t in a.blocks[0][0]] == \
        [(t.form, t.coef_kind, t.scale) for t in b.blocks[0][0]]
    assert np.array_equal(a.csc_structure()[1], b.csc_structure()[1])


def test_oracle_block_values_invariants():
    """The oracle's block assembly, pinned by what the mathematics fixes: a
    vector mass matrix sums to ncomp * area with every cross-component entry
    an explicit zero; a vector stiffness matrix annihilates any field that
    is constant per component; an SoA MultiFESpace of two scalar copies is
    the block-diagonal of the scalar matrix."""
    m = B.mixed_tri_quad_mesh(5, block=2, p_split=0.5, seed=7, backend=HOST)[0]
    om = _om(m)
    cp, rv, nz, n = O.assemble_blocks(om, 2, [2], {(0, 0): (["mass"], None)}, degree=4)
    assert np.isclose(nz.sum(), 2.0)
    U = B.TrialFESpace(LAG(2), m, size=2)
    comp = np.empty(n, dtype=np.int64)
    for k in range(2):
        comp[U.component_map[k]] = k
    col = np.repeat(np.arange(n), np.diff(cp))
    cross = comp[rv - 1] != comp[col]
    assert cross.sum() == (~cross).sum() and np.all(nz[cross] == 0.0)
    cp, rv, nz, n = O.assemble_blocks(om, 2, [2], {(0, 0): (["stiffness"], None)}, degree=4)
    import scipy.sparse as sp
    A = sp.csc_matrix((nz, rv - 1, cp - 1), shape=(n, n))
    x = np.where(comp == 0, 3.0, -1.5)
    assert np.abs(A @ x).max() <= 1e-12 * np.abs(nz).max()
    s_cp, s_rv, s_nz, ns = O.assemble_bilinear(om, 2, ["stiffness"], degree=4)
    S = sp.csc_matrix((s_nz, s_rv - 1, s_cp - 1), shape=(ns, ns))
    cp, rv, nz, n = O.assemble_blocks(om, 2, [1, 1], {(0, 0): (["stiffness"], None),
                                                      (1, 1): (["stiffness"], None)},
                                      aos=False, degree=4)
    A = sp.csc_matrix((nz, rv - 1, cp - 1), shape=(n, n))
    assert n == 2 * ns and (A - sp.block_diag([S, S])).count_nonzero() == 0


def _full_coupling(m, aos):
    sp = [B.TrialFESpace(LAG(2), m), B.TrialFESpace(LAG(2), m)]
    U = B.MultiFESpace(*sp, arrayOfStruct=aos)
    V = B.MultiFESpace(*[B.TestFESpace(s) for s in sp], arrayOfStruct=aos)
    dΩ = B.Measure(B.CellDomain(m), 4)

    def a(u, v):
        return B.integrate(u[0] * v[0] + B.dot(B.grad(u[1]), B.grad(v[1])) + u[0] * v[1]
                           + 3.0 * B.dot(B.grad(u[1]), B.grad(v[0])), dΩ)
    blocks = {(0, 0): (["mass"], None), (1, 1): (["stiffness"], None), (1, 0): (["mass"], None),
              (0, 1): (["stiffness"], np.full(m.ncells, 3.0))}
    return U, V, a, blocks


def test_full_coupling_and_many_fields_patterns():
    """Fully coupled MultiFESpaces (AoS, SoA) and 17 fields whose 33 coupled
    pairs share one scalar block: patterns bit-exact with the oracle."""
    m = _mesh("mixed")
    for aos in (True, False):
        U, V, a, blocks = _full_coupling(m, aos)
        plan = B.MultiBilinearPlan(a, U, V)
        ocp, orv, _, _ = O.assemble_blocks(_om(m), 2, [1, 1], blocks, aos=aos, degree=4)
        cp, rv = plan.csc_structure()
        assert np.array_equal(cp + 1, ocp) and np.array_equal(rv + 1, orv)
    U, V, a, blocks = _many_fields(m, 17)
    plan = B.MultiBilinearPlan(a, U, V)
    assert len(plan.blocks) == 1
    ocp, orv, _, _ = O.assemble_blocks(_om(m), 1, [1] * 17, blocks, degree=2)
    cp, rv = plan.csc_structure()
    assert np.array_equal(cp + 1, ocp) and np.array_equal(rv + 1, orv)


def _many_fields(m, nf):
    sp = [B.TrialFESpace(LAG(1), m) for _ in range(nf)]
    U = B.MultiFESpace(*sp)
    V = B.MultiFESpace(*[B.TestFESpace(s) for s in sp])
    dΩ = B.Measure(B.CellDomain(m), 2)

    def a(u, v):
        e = B.integrate(u[0] * v[0], dΩ)
        for i in range(1, nf):
            e = e + B.integrate(u[i] * v[i] + u[i - 1] * v[i], dΩ)
        return e
    blocks = {(0, 0): (["mass"], None)}
    for i in range(1, nf):
        blocks[(i, i)] = (["mass"], None)
        blocks[(i, i - 1)] = (["mass"], None)
    return U, V, a, blocks


def test_grad_scalar_dot_vector_lowering():
    """∇(p)⋅v over (v vector of the mesh's dimension, p scalar) lowers to one
    advection term per component of v (the form of test_assembler.jl:201-203);
    the terms a scalar test space would need from a vector trial space (q ∇⋅u)
    are refused."""
    from bcubehip.assembler import Fields, _integrals
    from bcubehip.forms import lower_multi_bilinear
    m = B.one_cell_mesh(":tri", backend=HOST)
    U_xy = B.TrialFESpace(LAG(1), m, size=2)
    U_z, U_p = B.TrialFESpace(LAG(1), m), B.TrialFESpace(LAG(1), m)
    U = B.MultiFESpace(U_p, U_xy, U_z)
    V = B.MultiFESpace(B.TestFESpace(U_p), B.TestFESpace(U_xy), B.TestFESpace(U_z))
    dΩ = B.Measure(B.CellDomain(m), 3)
    fu = Fields(U)

    def lower(a):
        items = _integrals(a(fu.placeholders(U, "trial"), fu.placeholders(V, "test")))
        return lower_multi_bilinear(items, fu.sizes, m.dim)
    out = lower(lambda u, v: B.integrate(B.dot(u[1], v[1]) + B.dot(u[0], v[2])
                                         + 2.0 * B.dot(B.grad(u[0]), v[1]), dΩ))
    assert out[(1, 1)] == {(0, 0): [("mass", "const", None, 1.0)],
                           (1, 1): [("mass", "const", None, 1.0)]}
    assert out[(2, 0)] == {(0, 0): [("mass", "const", None, 1.0)]}
    assert out[(1, 0)] == {(0, 0): [(("adv", 0), "const", None, 2.0)],
                           (1, 0): [(("adv", 1), "const", None, 2.0)]}
    with pytest.raises(NotImplementedError):
        lower(lambda u, v: B.integrate(B.dot(u[1], B.grad(v[0])), dΩ))
    with pytest.raises(NotImplementedError):
        lower(lambda u, v: B.integrate(u[0] * v[1], dΩ))
    # the global pattern of the coupled system: every (component of v, p) pair
    # of each cell, bit-exact with the oracle's block assembly
    plan = B.MultiBilinearPlan(lambda u, v: B.integrate(B.dot(u[1], v[1]) + B.dot(u[0], v[2])
                                                        + B.dot(B.grad(u[0]), v[1]), dΩ), U, V)
    ocp, orv, _, n = O.assemble_blocks(_om(m), 1, [1, 2, 1], {
        (1, 1): (["mass"], None), (2, 0): (["mass"], None), (1, 0): (["mass"], None)}, degree=3)
    cp, rv = plan.csc_structure()
    assert plan.nrows == n == 12
    assert np.array_equal(cp + 1, ocp) and np.array_equal(rv + 1, orv)
