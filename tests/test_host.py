"""Host-side logic of the product, checked on the CPU against the oracle.

The C-ABI library is loaded without a GPU context (host-only handles): mesh
generation, DofHandler numbering and the CSR pattern must be bit-identical to
the oracle's restatement of the reference; error paths must report cleanly.
"""
import ctypes as C
import json
import os
import re

import numpy as np
import pytest

import bcubehip as B
from bcubehip import _lib as L
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = B.BcubeBackendHostOnly()
CASES = json.load(open(os.path.join(ROOT, "tests", "golden", "dofhandler_cases.json")))


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "bcubehip.h")).read()
    declared = set(re.findall(r"^(?:int|const char\*)\s+(bh_\w+)\(", hdr, re.M))
    assert len(declared) >= 20
    lib = L.lib()
    for name in sorted(declared):
        assert hasattr(lib, name), name
    assert declared == set(L.EXPORTED)


@pytest.mark.parametrize("nx,ny,order", [(65, 65, 1), (3, 2, 1), (4, 5, 2)])
def test_rectangle_mesh_matches_oracle(nx, ny, order):
    m = B.rectangle_mesh(nx, ny, order=order, backend=HOST)
    o = O.rectangle_mesh(nx, ny, order=order)
    assert np.array_equal(m.coords, o.coords)
    assert np.array_equal(m.c2n, o.c2n)


def test_hexa_mesh_matches_oracle():
    m = B.hexa_mesh(6, 5, 4, -0.5, 0.5, 0, 2, 1, 3, backend=HOST)
    o = O.hexa_mesh(6, 5, 4, -0.5, 0.5, 0, 2, 1, 3)
    assert np.array_equal(m.coords, o.coords)
    assert np.array_equal(m.c2n, o.c2n)


def _mesh_from_case(m):
    if m["kind"] == "rectangle":
        return B.rectangle_mesh(m["nx"], m["ny"], order=m["order"], backend=HOST)
    nodes = np.array(m["nodes"], float)
    return B.Mesh(nodes, m["celltypes"], [np.array(c) - 1 for c in m["cells"]], HOST)


@pytest.mark.parametrize("case", CASES["cases"], ids=[c["name"] for c in CASES["cases"]])
def test_product_dofhandler_fixtures(case):
    """test/fespace/test_dofhandler.jl integer vectors through bh_dofhandler."""
    mesh = _mesh_from_case(case["mesh"])
    dhl = B.DofHandler(mesh, B.FunctionSpace("Lagrange", case["order"]), 1, case["continuous"])
    for k, want in case["expect"].items():
        assert dhl.get_dof(int(k) - 1).tolist() == want


@pytest.mark.parametrize("shape,n,order", [("hexa", 5, 1), ("hexa", 6, 2), ("hexa", 4, 3),
                                           ("quad", 9, 1), ("quad", 7, 2), ("quad", 6, 3),
                                           ("quad", 5, 4)])
def test_product_dofhandler_matches_oracle(shape, n, order):
    if shape == "hexa":
        m, o = B.hexa_mesh(n, n - 1, n + 1, backend=HOST), O.hexa_mesh(n, n - 1, n + 1)
    else:
        m, o = B.rectangle_mesh(n, n + 2, backend=HOST), O.rectangle_mesh(n, n + 2)
    for cont in (True, False):
        dhl = B.DofHandler(m, B.FunctionSpace("Lagrange", order), 1, cont)
        ig, off, nd = O.dofhandler(o, order, cont)
        assert dhl.ndofs_tot == nd
        assert np.array_equal(dhl.iglob, ig) and np.array_equal(dhl.offset, off)


def _plan(mesh, order, degree):
    U = B.TrialFESpace(B.FunctionSpace("Lagrange", order), mesh)
    V = B.TestFESpace(U)
    dΩ = B.Measure(B.CellDomain(mesh), degree)
    a = lambda u, v: B.integrate(B.dot(B.grad(u), B.grad(v)), dΩ) + B.integrate(u * v, dΩ)
    return B.BilinearPlan(a, U, V)


@pytest.mark.parametrize("cfg", ["C1 64x64 Q1", "C2 8^3 Q2", "hexa 5^3 Q1", "quad Q3"])
def test_csr_pattern_bit_exact(cfg):
    """The CSC structure sparse() returns (assembler.jl:185) equals the pattern."""
    if cfg == "C1 64x64 Q1":
        m, o, order, deg = (B.rectangle_mesh(65, 65, backend=HOST), O.rectangle_mesh(65, 65), 1, 2)
    elif cfg == "C2 8^3 Q2":
        m, o, order, deg = B.hexa_mesh(9, 9, 9, backend=HOST), O.hexa_mesh(9, 9, 9), 2, 5
    elif cfg == "hexa 5^3 Q1":
        m, o, order, deg = B.hexa_mesh(6, 6, 6, backend=HOST), O.hexa_mesh(6, 6, 6), 1, 2
    else:
        m, o, order, deg = B.rectangle_mesh(6, 5, backend=HOST), O.rectangle_mesh(6, 5), 3, 6
    plan = _plan(m, order, deg)
    colptr, rowval = plan.csc_structure()
    ocp, orv, _, n = O.assemble_bilinear(o, order, ["mass"], degree=deg)
    assert plan.nrows == plan.ncols == n
    assert np.array_equal(colptr + 1, ocp)
    assert np.array_equal(rowval + 1, orv)


def test_pattern_sizes_c2_formula():
    """Q2 on an n^3 hexa grid: the pattern is the tensor product of the 1D Q2
    pattern on 2n+1 nodes (5 entries per interior vertex row, 3 per boundary
    vertex row and per mid-edge row)."""
    n = 4
    plan = _plan(B.hexa_mesh(n + 1, n + 1, n + 1, backend=HOST), 2, 5)
    nodes1 = 2 * n + 1
    row1 = []
    for i in range(nodes1):
        if i % 2 == 1:
            row1.append(3)
        else:
            lo = max(0, i - 2)
            hi = min(nodes1 - 1, i + 2)
            row1.append(hi - lo + 1)
    assert plan.nnz == sum(row1) ** 3
    assert plan.max_row == 125


def test_error_paths_report_without_corrupting_state():
    lib = L.lib()
    coords = np.zeros(8)
    ct = np.zeros(1, dtype=np.int32)
    ptr = np.array([0, 4], dtype=np.int64)
    bad = np.array([0, 1, 2, 7], dtype=np.int64)  # node 7 of 4
    h = C.c_void_p()
    rc = lib.bh_mesh_create(None, 2, 4, coords.ctypes.data, 1, ct.ctypes.data, ptr.ctypes.data,
                            bad.ctypes.data, C.byref(h))
    assert rc == -5 and b"out of range" in lib.bh_last_error(None)
    rc = lib.bh_mesh_create(None, 2, 4, coords.ctypes.data, 1, ct.ctypes.data,
                            np.array([0, 3], dtype=np.int64).ctypes.data,
                            np.array([0, 1, 2], dtype=np.int64).ctypes.data, C.byref(h))
    assert rc == -1
    with pytest.raises(L.BcubeHipError):
        B.DofHandler(B.hexa_mesh(3, 3, 3, backend=HOST), B.FunctionSpace("Lagrange", 5))
    # a good mesh still works afterwards
    m = B.rectangle_mesh(3, 3, backend=HOST)
    assert B.DofHandler(m, B.FunctionSpace("Lagrange", 1)).ndofs_tot == 9


def test_unsupported_forms_raise():
    from bcubehip.forms import Placeholder, lower_bilinear, lower_linear
    u, v = Placeholder("trial", None), Placeholder("test", None)
    with pytest.raises(NotImplementedError):
        lower_bilinear(B.dot(B.grad(u), B.grad(u)))
    with pytest.raises(NotImplementedError):
        lower_linear(B.dot(B.grad(v), B.grad(v)))
    f, kind, data, scale = lower_bilinear(2.0 * B.dot(3.0 * B.grad(u), B.grad(v)))
    assert (f, kind, scale) == ("stiffness", "const", 6.0)
    f, kind, data, scale = lower_bilinear(u * 0.5 * v)
    assert (f, kind, scale) == ("mass", "const", 0.5)


def test_batched_meshes_are_block_diagonal_in_numbering_and_pattern():
    """Concatenated instances keep per-instance numbering (shifted) and the
    pattern is the block diagonal of the per-instance patterns."""
    sides = [3, 5, 4, 6]
    ms = [B.rectangle_mesh(s + 1, s + 1, backend=HOST) for s in sides]
    for order in (1, 2):
        big, coff, noff = B.batch_meshes(ms, backend=HOST)
        dbig = B.DofHandler(big, B.FunctionSpace("Lagrange", order))
        shift = 0
        plan_big = _plan(big, order, 2 * order)
        cb, rb = plan_big.csc_structure()
        for k, m in enumerate(ms):
            d = B.DofHandler(m, B.FunctionSpace("Lagrange", order))
            for c in range(m.ncells):
                assert np.array_equal(dbig.get_dof(coff[k] + c), d.get_dof(c) + shift)
            pk = _plan(m, order, 2 * order)
            ck, rk = pk.csc_structure()
            cols = slice(shift, shift + d.ndofs_tot)
            seg = slice(cb[shift], cb[shift + d.ndofs_tot])
            assert np.array_equal(cb[cols] - cb[shift], ck[:-1])
            assert np.array_equal(rb[seg] - shift, rk)
            shift += d.ndofs_tot
        assert shift == dbig.ndofs_tot
