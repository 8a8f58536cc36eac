"""ctypes bindings to the CPU oracle (``liboracle.so`` built from bcube_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()`` and
the CPU-baseline leg of ``bench.py`` -- never by the product package.  Every
function restates Bcube.jl's Float64 CPU assembly (file:line citations live in
bcube_oracle.cpp next to each routine).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

CT = {"Quad4": 0, "Quad9": 1, "Tri3": 2, "Hexa8": 3, "Hexa27": 4}
SHAPE = {"Square": 0, "Triangle": 1, "Cube": 2, "Line": -1}
TERM = {"stiffness": 0, "mass": 1}
SOURCE = {"one": 0, "sinsin": 1, "x": 2, "const": 3}

_d = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_l = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_i = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def build(force: bool = False) -> str:
    path = os.path.join(_HERE, "liboracle.so")
    src = os.path.join(_HERE, "bcube_oracle.cpp")
    if force or not os.path.exists(path) or os.path.getmtime(path) < os.path.getmtime(src):
        subprocess.run(["make", "-C", _HERE, "-B", "liboracle.so"], check=True,
                       stdout=subprocess.DEVNULL)
    return path


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        L.or_last_error.restype = C.c_char_p
        L.or_rectangle_mesh.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                        C.c_double, C.c_int, _d, _l]
        L.or_hexa_mesh.argtypes = [C.c_int, C.c_int, C.c_int] + [C.c_double] * 6 + [_d, _l]
        L.or_quadrature.argtypes = [C.c_int, C.c_int, C.c_int, _d, _d]
        L.or_quadrature.restype = C.c_int
        L.or_shape_functions.argtypes = [C.c_int, C.c_int, _d, _d, _d]
        L.or_shape_functions.restype = C.c_int
        L.or_geometry.argtypes = [C.c_int, _d, _d, _d, _d, _d, _d]
        L.or_geometry.restype = C.c_int
        L.or_densify.argtypes = [C.c_int64, _l, _l]
        mesh_args = [C.c_int, C.c_int64, _d, C.c_int64, _i, _l, _l]
        L.or_dofhandler.argtypes = mesh_args + [C.c_int, C.c_int, _l, _l]
        L.or_dofhandler.restype = C.c_int64
        L.or_dofhandler_n.argtypes = mesh_args + [C.c_int, C.c_int, C.c_int, _l, _l]
        L.or_dofhandler_n.restype = C.c_int64
        L.or_assemble_bilinear.argtypes = mesh_args + [C.c_int, C.c_int, C.c_int, _i,
                                                       C.c_void_p, C.c_int]
        L.or_assemble_bilinear.restype = C.c_void_p
        L.or_assemble_linear.argtypes = mesh_args + [C.c_int, C.c_int, C.c_int, C.c_double,
                                                     C.c_int]
        L.or_assemble_linear.restype = C.c_void_p
        L.or_assemble_bilinear_rect.argtypes = mesh_args + [C.c_int, C.c_int, C.c_int, C.c_int,
                                                            _i, C.c_void_p, C.c_int]
        L.or_assemble_bilinear_rect.restype = C.c_void_p
        L.or_assemble_linear_fe.argtypes = mesh_args + [C.c_int, C.c_int, C.c_int, _d, C.c_int,
                                                        _d, C.c_int]
        L.or_assemble_linear_fe.restype = C.c_void_p
        for n in ("or_result_ndofs", "or_result_nnz", "or_result_nrows"):
            getattr(L, n).argtypes = [C.c_void_p]
            getattr(L, n).restype = C.c_int64
        L.or_result_csc.argtypes = [C.c_void_p, _l, _l, _d]
        L.or_result_vector.argtypes = [C.c_void_p, _d]
        L.or_result_dofs.argtypes = [C.c_void_p, _l, _l]
        L.or_result_free.argtypes = [C.c_void_p]
        L.or_dg_transport.argtypes = mesh_args + [C.c_int, _d, C.c_int, C.c_int, C.c_int, _d, _d]
        L.or_dg_transport.restype = C.c_int64
        L.or_geometry27.argtypes = [_d, _d, _d, _d, _d, _d]
        L.or_geometry27.restype = C.c_int
        L.or_bilinear_timed.argtypes = mesh_args + [C.c_int, C.c_int, C.c_int, _i, C.c_void_p,
                                                    C.c_int, C.c_int, C.c_int, _d]
        L.or_bilinear_timed.restype = C.c_void_p
        L.or_linear_timed.argtypes = mesh_args + [C.c_int, C.c_int, C.c_int, C.c_double, C.c_int,
                                                  C.c_int, C.c_int, _d]
        L.or_linear_timed.restype = C.c_void_p
        L.or_dg_timed.argtypes = mesh_args + [C.c_int, _d, C.c_int, C.c_int, C.c_int, _d, _d,
                                              C.c_int, C.c_int, _d]
        L.or_dg_timed.restype = C.c_int64
        L.or_order_space.argtypes = mesh_args + [_i, C.c_int, _l, _i]
        L.or_order_space.restype = C.c_int64
        L.or_assemble_bilinear_orders.argtypes = mesh_args + [_i, C.c_int, _i, C.c_void_p, _i]
        L.or_assemble_bilinear_orders.restype = C.c_void_p
        L.or_assemble_linear_orders.argtypes = mesh_args + [_i, C.c_int, C.c_double, _i]
        L.or_assemble_linear_orders.restype = C.c_void_p
        L.or_bilinear_orders_timed.argtypes = mesh_args + [_i, C.c_int, _i, C.c_void_p, _i,
                                                           C.c_int, C.c_int, _d]
        L.or_bilinear_orders_timed.restype = C.c_void_p
        _LIB = L
    return _LIB


def _err():
    return lib().or_last_error().decode()


class OMesh:
    """Plain mesh arrays as the oracle consumes them (0-based c2n CSR)."""

    def __init__(self, dim, coords, celltypes, c2n_ptr, c2n):
        self.dim = dim
        self.coords = np.ascontiguousarray(coords, dtype=np.float64).reshape(-1)
        self.celltypes = np.ascontiguousarray(celltypes, dtype=np.int32)
        self.c2n_ptr = np.ascontiguousarray(c2n_ptr, dtype=np.int64)
        self.c2n = np.ascontiguousarray(c2n, dtype=np.int64)

    @property
    def nnodes(self):
        return self.coords.size // self.dim

    @property
    def ncells(self):
        return self.celltypes.size

    def args(self):
        return (self.dim, self.nnodes, self.coords, self.ncells, self.celltypes,
                self.c2n_ptr, self.c2n)

    @staticmethod
    def uniform(dim, coords, ctype, c2n_2d):
        c2n_2d = np.asarray(c2n_2d, dtype=np.int64)
        nc, k = c2n_2d.shape
        return OMesh(dim, coords, np.full(nc, ctype, np.int32),
                     np.arange(nc + 1, dtype=np.int64) * k, c2n_2d.reshape(-1))


def rectangle_mesh(nx, ny, xmin=0.0, xmax=1.0, ymin=0.0, ymax=1.0, order=1):
    """mesh_generator.jl:218-434 (Quad4 for order 1, Quad9 for order 2)."""
    if order == 1:
        nn, k, ct = nx * ny, 4, CT["Quad4"]
    else:
        nn = nx * ny + (nx - 1) * ny + (ny - 1) * nx + (nx - 1) * (ny - 1)
        k, ct = 9, CT["Quad9"]
    coords = np.zeros(nn * 2)
    c2n = np.zeros((nx - 1) * (ny - 1) * k, dtype=np.int64)
    lib().or_rectangle_mesh(nx, ny, xmin, xmax, ymin, ymax, order, coords, c2n)
    return OMesh.uniform(2, coords, ct, c2n.reshape(-1, k))


def hexa_mesh(nx, ny, nz, xmin=0.0, xmax=1.0, ymin=0.0, ymax=1.0, zmin=0.0, zmax=1.0):
    """mesh_generator.jl:453-552."""
    coords = np.zeros(nx * ny * nz * 3)
    c2n = np.zeros((nx - 1) * (ny - 1) * (nz - 1) * 8, dtype=np.int64)
    lib().or_hexa_mesh(nx, ny, nz, xmin, xmax, ymin, ymax, zmin, zmax, coords, c2n)
    return OMesh.uniform(3, coords, CT["Hexa8"], c2n.reshape(-1, 8))


def quadrature(shape, degree, qtype=0):
    s = SHAPE[shape]
    w = np.zeros(4096)
    x = np.zeros(4096 * 3)
    n = lib().or_quadrature(s, degree, qtype, w, x)
    if n < 0:
        raise RuntimeError(_err())
    d = 1 if s < 0 else (3 if s == 2 else 2)
    return w[:n].copy(), x[: n * d].reshape(n, d).copy()


def shape_functions(shape, order, xi):
    s = SHAPE[shape]
    d = 3 if s == 2 else 2
    n = (order + 1) * (order + 2) // 2 if s == 1 else (order + 1) ** d
    val = np.zeros(n)
    grad = np.zeros(n * d)
    r = lib().or_shape_functions(s, order, np.ascontiguousarray(xi, dtype=np.float64), val, grad)
    if r < 0:
        raise RuntimeError(_err())
    return val, grad.reshape(n, d)


def geometry(celltype, cnodes, xi):
    ct = CT[celltype]
    d = 3 if ct in (3, 4) else 2
    J = np.zeros(d * d)
    Ji = np.zeros(d * d)
    det = np.zeros(1)
    x = np.zeros(d)
    r = lib().or_geometry(ct, np.ascontiguousarray(cnodes, dtype=np.float64).reshape(-1),
                          np.ascontiguousarray(xi, dtype=np.float64), J, Ji, det, x)
    if r < 0:
        raise RuntimeError(_err())
    return J.reshape(d, d), Ji.reshape(d, d), det[0], x


def densify(a):
    a = np.ascontiguousarray(a, dtype=np.int64)
    out = np.zeros_like(a)
    lib().or_densify(a.size, a, out)
    return out


def dofhandler(mesh: OMesh, order, continuous=True):
    """Returns (iglob 1-based, offset, ndofs_tot)."""
    nd = {0: (order + 1) ** 2, 1: (order + 1) ** 2, 2: (order + 1) * (order + 2) // 2,
          3: (order + 1) ** 3, 4: (order + 1) ** 3}
    total = int(sum(nd[int(t)] for t in mesh.celltypes))
    iglob = np.zeros(total, dtype=np.int64)
    offset = np.zeros(mesh.ncells, dtype=np.int64)
    n = lib().or_dofhandler(*mesh.args(), order, int(continuous), iglob, offset)
    if n < 0:
        raise RuntimeError(_err())
    return iglob, offset, int(n)


def dofhandler_n(mesh: OMesh, order, ncomp, continuous=True):
    """DofHandler(mesh, fs, ncomp, continuous) (dofhandler.jl:37-156): iglob
    (1-based), offset (ncells, ncomp) 0-based, ndofs_tot."""
    nd = {0: (order + 1) ** 2, 1: (order + 1) ** 2, 2: (order + 1) * (order + 2) // 2,
          3: (order + 1) ** 3, 4: (order + 1) ** 3}
    total = int(sum(nd[int(t)] for t in mesh.celltypes)) * ncomp
    iglob = np.zeros(total, dtype=np.int64)
    offset = np.zeros(len(mesh.celltypes) * ncomp, dtype=np.int64)
    n = lib().or_dofhandler_n(*mesh.args(), order, int(continuous), ncomp, iglob, offset)
    if n < 0:
        raise RuntimeError(_err())
    return iglob, offset.reshape(-1, ncomp), int(n)


def cell_dofs(iglob, offset, ncomp, ndofs_cell, icell):
    """get_dofs(U, icell) (dofhandler.jl:489-505): 1-based, components in turn."""
    o = offset[icell, 0]
    return iglob[o:o + ncomp * ndofs_cell[icell]]


def multi_mapping(cell_dofs_of, ndofs, ncells, aos=True):
    """_build_mapping_AoS / _build_mapping_SoA (fespace.jl:604-640), as loops:
    cell_dofs_of[i](icell) -> 1-based dofs of space i; returns (ndofs_tot,
    mappings with mapping[i][d-1] = global 1-based id)."""
    mapping = [np.zeros(n, dtype=np.int64) for n in ndofs]
    tot = 0
    order = ([(c, i) for c in range(ncells) for i in range(len(ndofs))] if aos
             else [(c, i) for i in range(len(ndofs)) for c in range(ncells)])
    for c, i in order:
        for d in cell_dofs_of[i](c):
            if mapping[i][d - 1] == 0:
                tot += 1
                mapping[i][d - 1] = tot
    assert all(np.all(m != 0) for m in mapping), "invalid mapping"
    return tot, mapping


def assemble_blocks(mesh: OMesh, order, sizes, blocks, aos=True, degree=2, continuous=True):
    """assemble_bilinear over a MultiFESpace of spaces of sizes[i] components
    (size 1: scalar) on one mesh and order (assembler.jl:327-420):
    blocks[(i, j)] = (terms, eta) couples test space i with trial space j; a
    pair of vector spaces carries the component-diagonal forms ∫c u⋅v,
    ∫η ∇u ⊡ ∇v.  The COO of each pair (every component pair of its cells'
    element matrices, zero or not) is mapped through the spaces' numberings
    and summed by `sparse`.  Returns (colptr, rowval 1-based, nzval, n)."""
    om = mesh
    ncells = len(om.celltypes)
    sig, soff, _ = dofhandler(om, order, continuous)
    nd = np.diff(np.append(soff, sig.size))
    vec, dofs_of, ndofs = [], [], []
    for ncomp in sizes:
        ig, off, n = dofhandler_n(om, order, ncomp, continuous)
        # component k's copy of each scalar dof, read off the cells
        table = np.full((ncomp, int(sig.max())), -1, dtype=np.int64)
        for c in range(ncells):
            g = sig[soff[c]:soff[c] + nd[c]] - 1
            for k in range(ncomp):
                v = ig[off[c, k]:off[c, k] + nd[c]] - 1
                assert np.all((table[k, g] == -1) | (table[k, g] == v)), "inconsistent numbering"
                table[k, g] = v
        vec.append(table)
        dofs_of.append(lambda c, ig=ig, off=off, ncomp=ncomp: cell_dofs(ig, off, ncomp, nd, c))
        ndofs.append(n)
    ntot, mapping = multi_mapping(dofs_of, ndofs, ncells, aos)
    rows, cols, vals = [], [], []
    for (i, j), (terms, eta) in blocks.items():
        cp, rv, nz, _ = assemble_bilinear(om, order, terms, eta=eta, degree=degree,
                                          continuous=continuous)
        hcol = np.repeat(np.arange(cp.size - 1), np.diff(cp))  # trial (column) scalar dof
        grow = rv - 1                                           # test (row) scalar dof
        for ki in range(sizes[i]):
            for kj in range(sizes[j]):
                diag = sizes[i] == 1 or ki == kj
                rows.append(mapping[i][vec[i][ki, grow]] - 1)
                cols.append(mapping[j][vec[j][kj, hcol]] - 1)
                vals.append(nz if diag else np.zeros_like(nz))
    r, c, v = np.concatenate(rows), np.concatenate(cols), np.concatenate(vals)
    o = np.lexsort((r, c))
    r, c, v = r[o], c[o], v[o]
    new = np.ones(r.size, dtype=bool)
    new[1:] = (r[1:] != r[:-1]) | (c[1:] != c[:-1])
    start = np.flatnonzero(new)
    nzval = np.add.reduceat(v, start) if v.size else v
    r, c = r[start], c[start]
    colptr = np.concatenate([[0], np.cumsum(np.bincount(c, minlength=ntot))]) + 1
    return colptr, r + 1, nzval, ntot


def assemble_bilinear(mesh: OMesh, order, terms, eta=None, degree=2, continuous=True):
    """CSC (1-based colptr/rowval, nzval) of Σ terms, in SparseMatrixCSC layout."""
    t = np.ascontiguousarray([TERM[x] for x in terms], dtype=np.int32)
    eta_p = None
    if eta is not None:
        eta = np.ascontiguousarray(eta, dtype=np.float64)
        eta_p = eta.ctypes.data
    h = lib().or_assemble_bilinear(*mesh.args(), order, int(continuous), len(t), t, eta_p, degree)
    if not h:
        raise RuntimeError(_err())
    try:
        n = lib().or_result_ndofs(h)
        nnz = lib().or_result_nnz(h)
        colptr = np.zeros(n + 1, dtype=np.int64)
        rowval = np.zeros(nnz, dtype=np.int64)
        nzval = np.zeros(nnz)
        lib().or_result_csc(h, colptr, rowval, nzval)
    finally:
        lib().or_result_free(h)
    return colptr, rowval, nzval, int(n)


def assemble_bilinear_rect(mesh: OMesh, order_u, order_v, terms, eta=None, degree=2,
                           continuous=True):
    """assemble_bilinear(a, U, V) with U of order_u and V of order_v on one mesh:
    CSC (1-based colptr over U's dofs, rowval over V's), nzval, (nrows, ncols)."""
    t = np.ascontiguousarray([TERM[x] for x in terms], dtype=np.int32)
    eta_p = None
    if eta is not None:
        eta = np.ascontiguousarray(eta, dtype=np.float64)
        eta_p = eta.ctypes.data
    h = lib().or_assemble_bilinear_rect(*mesh.args(), order_u, order_v, int(continuous), len(t),
                                        t, eta_p, degree)
    if not h:
        raise RuntimeError(_err())
    try:
        n = lib().or_result_ndofs(h)
        m = lib().or_result_nrows(h)
        nnz = lib().or_result_nnz(h)
        colptr = np.zeros(n + 1, dtype=np.int64)
        rowval = np.zeros(nnz, dtype=np.int64)
        nzval = np.zeros(nnz)
        lib().or_result_csc(h, colptr, rowval, nzval)
    finally:
        lib().or_result_free(h)
    return colptr, rowval, nzval, (int(m), int(n))


def assemble_linear_fe(mesh: OMesh, order_u, u, order_v, kind="uv", c=None, degree=2,
                       continuous=True):
    """Linear forms with an FE function u (dof values of the order_u space),
    tested on the order_v space: kind "uv" ∫(u v)dΩ, "cu_grad" ∫((c u)⋅∇v)dΩ."""
    u = np.ascontiguousarray(u, dtype=np.float64)
    cv = np.ascontiguousarray(np.zeros(3) if c is None else np.resize(np.asarray(c, float), 3))
    h = lib().or_assemble_linear_fe(*mesh.args(), order_u, order_v, int(continuous), u,
                                    {"uv": 0, "cu_grad": 1}[kind], cv, degree)
    if not h:
        raise RuntimeError(_err())
    try:
        b = np.zeros(lib().or_result_ndofs(h))
        lib().or_result_vector(h, b)
    finally:
        lib().or_result_free(h)
    return b


def assemble_linear(mesh: OMesh, order, source="sinsin", const=0.0, degree=2, continuous=True):
    h = lib().or_assemble_linear(*mesh.args(), order, int(continuous), SOURCE[source],
                                 float(const), degree)
    if not h:
        raise RuntimeError(_err())
    try:
        n = lib().or_result_ndofs(h)
        b = np.zeros(n)
        lib().or_result_vector(h, b)
    finally:
        lib().or_result_free(h)
    return b


def _orders_args(mesh: OMesh, order, degree):
    order = np.ascontiguousarray(order, dtype=np.int32)
    assert order.size == mesh.ncells
    deg = np.zeros(5, dtype=np.int32)
    for p in range(1, 5):
        deg[p] = degree(p) if callable(degree) else int(degree)
    return order, deg


def order_slots(mesh: OMesh, order):
    """Slots per cell of a per-cell-order space: the largest cell's Lagrange
    dof count."""
    nd = {1: lambda p: (p + 1) ** 2, 2: lambda p: (p + 1) * (p + 2) // 2}
    return max(nd[1 if int(t) in (0, 1) else 2](int(p)) for t, p in zip(mesh.celltypes, order))


def order_space(mesh: OMesh, order):
    """Continuous Lagrange space of per-cell order on a 2D Quad4/Tri3 mesh
    (edges take their cells' lowest order; see bcube_oracle.cpp).  Returns
    (dofs (ncells, nl) 0-based with -1 on unused slots, eord (ncells, 4): the
    order of each local edge of lower order than its cell, 0 if free, ndofs)."""
    order = np.ascontiguousarray(order, dtype=np.int32)
    nl = order_slots(mesh, order)
    dofs = np.zeros(mesh.ncells * nl, dtype=np.int64)
    eord = np.zeros(mesh.ncells * 4, dtype=np.int32)
    n = lib().or_order_space(*mesh.args(), order, nl, dofs, eord)
    if n < 0:
        raise RuntimeError(_err())
    return dofs.reshape(-1, nl), eord.reshape(-1, 4), int(n)


def assemble_bilinear_orders(mesh: OMesh, order, terms, eta=None, degree=lambda p: max(2 * p, 4)):
    """CSC (1-based) of Σ terms on the per-cell-order space; degree: the
    Measure degree of the cells of order p (a function of p, or one int)."""
    order, deg = _orders_args(mesh, order, degree)
    t = np.ascontiguousarray([TERM[x] for x in terms], dtype=np.int32)
    eta_p = None
    if eta is not None:
        eta = np.ascontiguousarray(eta, dtype=np.float64)
        eta_p = eta.ctypes.data
    h = lib().or_assemble_bilinear_orders(*mesh.args(), order, len(t), t, eta_p, deg)
    if not h:
        raise RuntimeError(_err())
    try:
        n = lib().or_result_ndofs(h)
        nnz = lib().or_result_nnz(h)
        colptr = np.zeros(n + 1, dtype=np.int64)
        rowval = np.zeros(nnz, dtype=np.int64)
        nzval = np.zeros(nnz)
        lib().or_result_csc(h, colptr, rowval, nzval)
    finally:
        lib().or_result_free(h)
    return colptr, rowval, nzval, int(n)


def assemble_bilinear_orders_timed(mesh: OMesh, order, terms, eta=None,
                                   degree=lambda p: max(2 * p, 4), threads=1, reps=5):
    """assemble_bilinear_orders timed: the numbering once (untimed), then
    reps + 1 runs on `threads` threads (the first a warm-up).  Returns
    ((colptr, rowval, nzval, n) of the last run, t (reps, 2): COO build and
    sparse() seconds per timed run)."""
    order, deg = _orders_args(mesh, order, degree)
    tt = np.ascontiguousarray([TERM[x] for x in terms], dtype=np.int32)
    eta_p = None
    if eta is not None:
        eta = np.ascontiguousarray(eta, dtype=np.float64)
        eta_p = eta.ctypes.data
    t = np.zeros(2 * (reps + 1))
    h = lib().or_bilinear_orders_timed(*mesh.args(), order, len(tt), tt, eta_p, deg, int(threads),
                                       int(reps), t)
    if not h:
        raise RuntimeError(_err())
    try:
        n = lib().or_result_ndofs(h)
        nnz = lib().or_result_nnz(h)
        colptr = np.zeros(n + 1, dtype=np.int64)
        rowval = np.zeros(nnz, dtype=np.int64)
        nzval = np.zeros(nnz)
        lib().or_result_csc(h, colptr, rowval, nzval)
    finally:
        lib().or_result_free(h)
    return (colptr, rowval, nzval, int(n)), t.reshape(-1, 2)[1:]


def assemble_linear_orders(mesh: OMesh, order, source="one", const=0.0,
                           degree=lambda p: max(2 * p, 4)):
    order, deg = _orders_args(mesh, order, degree)
    h = lib().or_assemble_linear_orders(*mesh.args(), order, SOURCE[source], float(const), deg)
    if not h:
        raise RuntimeError(_err())
    try:
        n = lib().or_result_ndofs(h)
        b = np.zeros(n)
        lib().or_result_vector(h, b)
    finally:
        lib().or_result_free(h)
    return b


def dg_transport(mesh: OMesh, order, c, u, vol_degree, face_degree=None):
    """DG upwind linear-transport residual, r = l(v) for
    l(v) = int((c*u).grad(v))dOmega - int(upwind flux * jump(v))dGamma over the interior
    faces.  u: (ndofs, nfields) nodal values of the discontinuous Lagrange space
    (cell-by-cell numbering); returns r with the same shape."""
    u = np.ascontiguousarray(u, dtype=np.float64)
    if u.ndim == 1:
        u = u[:, None]
    nf = u.shape[1]
    r = np.zeros_like(u)
    c = np.ascontiguousarray(c, dtype=np.float64)
    fd = vol_degree if face_degree is None else face_degree
    n = lib().or_dg_transport(*mesh.args(), order, c, vol_degree, fd, nf, u, r)
    if n < 0:
        raise RuntimeError(_err())
    if n != u.shape[0]:
        raise ValueError(f"u has {u.shape[0]} rows, the space has {n} dofs")
    return r


def host_threads() -> int:
    """Host threads for the all-core CPU baseline: OMP_NUM_THREADS when set,
    else the cores this process may run on."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env)
    return len(os.sched_getaffinity(0))


def _csc_of(h):
    try:
        n = lib().or_result_ndofs(h)
        nnz = lib().or_result_nnz(h)
        colptr = np.zeros(n + 1, dtype=np.int64)
        rowval = np.zeros(nnz, dtype=np.int64)
        nzval = np.zeros(nnz)
        lib().or_result_csc(h, colptr, rowval, nzval)
    finally:
        lib().or_result_free(h)
    return colptr, rowval, nzval, int(n)


def assemble_bilinear_timed(mesh: OMesh, order, terms, eta=None, degree=2, continuous=True,
                            threads=1, reps=5):
    """assemble_bilinear timed: the numbering once (untimed), then reps + 1
    assemblies on `threads` threads, the first a warm-up.  Returns the CSC of
    the last run and the per-run seconds of the COO build and of sparse(),
    shape (reps, 2), warm-up dropped.  Any thread count gives the same CSC."""
    t = np.ascontiguousarray([TERM[x] for x in terms], dtype=np.int32)
    eta_p = None
    if eta is not None:
        eta = np.ascontiguousarray(eta, dtype=np.float64)
        eta_p = eta.ctypes.data
    times = np.zeros(2 * (reps + 1))
    h = lib().or_bilinear_timed(*mesh.args(), order, int(continuous), len(t), t, eta_p, degree,
                                int(threads), int(reps), times)
    if not h:
        raise RuntimeError(_err())
    return _csc_of(h), times.reshape(reps + 1, 2)[1:]


def assemble_linear_timed(mesh: OMesh, order, source="sinsin", const=0.0, degree=2,
                          continuous=True, threads=1, reps=5):
    """assemble_linear timed as assemble_bilinear_timed: (b, seconds per run)."""
    times = np.zeros(reps + 1)
    h = lib().or_linear_timed(*mesh.args(), order, int(continuous), SOURCE[source], float(const),
                              degree, int(threads), int(reps), times)
    if not h:
        raise RuntimeError(_err())
    try:
        b = np.zeros(lib().or_result_ndofs(h))
        lib().or_result_vector(h, b)
    finally:
        lib().or_result_free(h)
    return b, times[1:]


def dg_transport_timed(mesh: OMesh, order, c, u, vol_degree, face_degree=None, threads=1,
                       reps=5):
    """dg_transport timed (numbering and faces built once, untimed): (r, seconds per run)."""
    u = np.ascontiguousarray(u, dtype=np.float64)
    if u.ndim == 1:
        u = u[:, None]
    r = np.zeros_like(u)
    c = np.ascontiguousarray(c, dtype=np.float64)
    fd = vol_degree if face_degree is None else face_degree
    times = np.zeros(reps + 1)
    n = lib().or_dg_timed(*mesh.args(), order, c, vol_degree, fd, u.shape[1], u, r, int(threads),
                          int(reps), times)
    if n < 0:
        raise RuntimeError(_err())
    if n != u.shape[0]:
        raise ValueError(f"u has {u.shape[0]} rows, the space has {n} dofs")
    return r, times[1:]


def geometry27(cnodes, xi):
    """Hexa27 map x(xi), Jacobian J, J^-1 and |det J| (ForwardDiff-style)."""
    X = np.ascontiguousarray(cnodes, dtype=np.float64).reshape(27, 3)
    xi = np.ascontiguousarray(xi, dtype=np.float64)
    J = np.zeros(9); Ji = np.zeros(9); d = np.zeros(1); x = np.zeros(3)
    lib().or_geometry27(X, xi, J, Ji, d, x)
    return x, J.reshape(3, 3), Ji.reshape(3, 3), float(d[0])


def csc_to_scipy(colptr, rowval, nzval, n):
    import scipy.sparse as sp
    return sp.csc_matrix((nzval, rowval - 1, colptr - 1), shape=(n, n))


# ---------------------------------------------------------------------------
# Symmetric interior-penalty DG on Quad4 meshes: the reference's Poisson DG
# benchmark (benchmark/bench_assemble.jl:33-96).  Plain loops over faces and
# points, for small meshes.
# ---------------------------------------------------------------------------
_SQ_COORDS = np.array([[-1.0, -1.0], [1.0, -1.0], [1.0, 1.0], [-1.0, 1.0]])  # shape.jl:208
_SQ_F2N = [(0, 1), (1, 2), (2, 3), (3, 0)]                                 # shape.jl:254
_SQ_NORMALS = np.array([[0.0, -1.0], [1.0, 0.0], [0.0, 1.0], [-1.0, 0.0]])  # shape.jl:278


def quad_faces(mesh: OMesh):
    """_build_faces! (mesh.jl:328-374) on a Quad4 mesh: faces in creation
    order (each cell's sides in faces2nodes order), [cell_n, side_n, cell_p,
    side_p] with cell_p = -1 on the boundary."""
    faces, index = [], {}
    for c in range(len(mesh.celltypes)):
        nodes = mesh.c2n[mesh.c2n_ptr[c]:mesh.c2n_ptr[c + 1]]
        for s, (a, b) in enumerate(_SQ_F2N):
            key = tuple(sorted((int(nodes[a]), int(nodes[b]))))
            if key in index:
                faces[index[key]][2:] = [c, s]
            else:
                index[key] = len(faces)
                faces.append([c, s, -1, -1])
    return faces


def _face_to_cell(local_pair, xi):
    """mapping_face (mapping.jl:324-336): the Bar2 map of the face's two
    reference-square corners."""
    a, b = local_pair
    return 0.5 * (1.0 - xi) * _SQ_COORDS[a] + 0.5 * (1.0 + xi) * _SQ_COORDS[b]


def _side_eval(mesh, c, order, xi_cell, side):
    nodes = mesh.c2n[mesh.c2n_ptr[c]:mesh.c2n_ptr[c + 1]]
    X = mesh.coords.reshape(-1, 2)[nodes]
    J, Ji, det, x = geometry("Quad4", X, xi_cell)
    val, gref = shape_functions("Square", order, xi_cell)
    grad = gref @ Ji                                   # ∇φ = J⁻ᵀ ∇̂φ
    n = _SQ_NORMALS[side] @ Ji                         # normal: J⁻ᵀ n_ref, normalised
    return val, grad, n / np.linalg.norm(n), x


def sipg_quad(mesh: OMesh, order, degree, penalty, f=None, g=None, terms=("vol", "int", "bnd")):
    """A (scipy CSC) and b of a_Ω + a_Γ + a_Γb and l_Ω + l_Γb, with the
    discontinuous Lagrange numbering (cell by cell), restating:
      a_Ω(u,v)  = ∫ ∇v⋅∇u dΩ
      a_Γ(u,v)  = ∫ -jump(v,n)⋅avg(∇u) - avg(∇v)⋅jump(u,n) + penalty jump(v,n)⋅jump(u,n) dΓ
      a_Γb(u,v) = ∫ -v (∇u⋅n) - (∇v⋅n) u + penalty v u dΓb
      l_Ω(v)    = ∫ v f dΩ,   l_Γb(v) = ∫ -(∇v⋅n) g + penalty v g dΓb
    with jump(a, n) = a⁺ n⁺ + a⁻ n⁻ (cellfunction.jl:310), avg(a) = (a⁺ + a⁻)/2
    (bench_assemble.jl:31), each side's normal J⁻ᵀ n_ref of its own cell
    (mapping.jl:843-855), side⁺ points through faces2nodes(side_p)[indexin(f2n_p,
    f2n_n)] (eval_point.jl:164-206), and the face metric |x_b - x_a|/2 of the
    Bar2 face.  f, g: functions of one point (2,) -> float."""
    import scipy.sparse as sp

    nc = len(mesh.celltypes)
    nl = (order + 1) ** 2
    n = nc * nl
    I, J_, X = [], [], []
    b = np.zeros(n)
    if "vol" in terms:
        cp, rv, nz, _ = assemble_bilinear(mesh, order, ["stiffness"], degree=degree, continuous=False)
        col = np.repeat(np.arange(n), np.diff(cp))
        I.append(rv - 1); J_.append(col); X.append(nz)
    wq, xq = quadrature("Line", degree)
    xq = xq.reshape(-1)
    faces = quad_faces(mesh)
    coords = mesh.coords.reshape(-1, 2)
    for cn, sn, cpp, spp in faces:
        nodes_n = mesh.c2n[mesh.c2n_ptr[cn]:mesh.c2n_ptr[cn + 1]]
        fa, fb = _SQ_F2N[sn]
        dS = np.linalg.norm(coords[nodes_n[fb]] - coords[nodes_n[fa]]) / 2
        if cpp >= 0 and "int" in terms:
            nodes_p = mesh.c2n[mesh.c2n_ptr[cpp]:mesh.c2n_ptr[cpp + 1]]
            gn = [int(nodes_n[k]) for k in _SQ_F2N[sn]]
            gp = [int(nodes_p[k]) for k in _SQ_F2N[spp]]
            perm = [gn.index(v) for v in gp]                   # indexin(f2n_p, f2n_n)
            loc_p = tuple(_SQ_F2N[spp][k] for k in perm)
            blocks = {}
            for w, xi in zip(wq, xq):
                sides = {0: (cn, _side_eval(mesh, cn, order, _face_to_cell(_SQ_F2N[sn], xi), sn)),
                         1: (cpp, _side_eval(mesh, cpp, order, _face_to_cell(loc_p, xi), spp))}
                for t in (0, 1):
                    vt, gt, nt, _ = sides[t][1]
                    for r in (0, 1):
                        vr, gr, nr, _ = sides[r][1]
                        m = (-0.5 * np.outer(vt, gr @ nt) - 0.5 * np.outer(gt @ nr, vr)
                             + penalty * (nt @ nr) * np.outer(vt, vr))
                        blocks[(t, r)] = blocks.get((t, r), 0.0) + w * dS * m
            for (t, r), m in blocks.items():
                ct, cr = (cn, cpp)[t], (cn, cpp)[r]
                I.append(np.repeat(ct * nl + np.arange(nl), nl))
                J_.append(np.tile(cr * nl + np.arange(nl), nl))
                X.append(m.reshape(-1))
        if cpp < 0 and "bnd" in terms:
            m = 0.0
            bl = 0.0
            for w, xi in zip(wq, xq):
                v, gv, nn, x = _side_eval(mesh, cn, order, _face_to_cell(_SQ_F2N[sn], xi), sn)
                dn = gv @ nn
                m = m + w * dS * (-np.outer(v, dn) - np.outer(dn, v) + penalty * np.outer(v, v))
                if g is not None:
                    gx = g(x)
                    bl = bl + w * dS * (-dn * gx + penalty * v * gx)
            I.append(np.repeat(cn * nl + np.arange(nl), nl))
            J_.append(np.tile(cn * nl + np.arange(nl), nl))
            X.append(np.asarray(m).reshape(-1))
            if g is not None:
                b[cn * nl:(cn + 1) * nl] += bl
    if f is not None:
        wv, xv = quadrature("Square", degree)
        for c in range(nc):
            nodes = mesh.c2n[mesh.c2n_ptr[c]:mesh.c2n_ptr[c + 1]]
            Xc = coords[nodes]
            acc = np.zeros(nl)
            for w, xi in zip(wv, xv):
                _, _, det, x = geometry("Quad4", Xc, xi)
                val, _ = shape_functions("Square", order, xi)
                acc += w * det * f(x) * val
            b[c * nl:(c + 1) * nl] += acc
    A = sp.coo_matrix((np.concatenate(X), (np.concatenate(I), np.concatenate(J_))),
                      shape=(n, n)).tocsc()
    A.sum_duplicates()
    A.sort_indices()
    return A, b


def dg_node_coords(mesh: OMesh, order):
    """Physical coordinates of the Lagrange nodes of each cell (uniform nodes,
    first index fastest), cell by cell: the discontinuous dofs' points."""
    t = np.linspace(-1.0, 1.0, order + 1)
    pts = np.array([[t[i], t[j]] for j in range(order + 1) for i in range(order + 1)])
    out = []
    coords = mesh.coords.reshape(-1, 2)
    for c in range(len(mesh.celltypes)):
        nodes = mesh.c2n[mesh.c2n_ptr[c]:mesh.c2n_ptr[c + 1]]
        for p in pts:
            out.append(geometry("Quad4", coords[nodes], p)[3])
    return np.array(out)


# ---------------------------------------------------------------------------
# Euler DG residual of the reference's COVO benchmark (benchmark/covo.jl):
# volume flux (_flux_Ω, covo.jl:44-58), Roe flux on interior and periodic
# faces (flux_roe, covo.jl:69-150), pressure (covo.jl:185-190).  Fields are
# the five scalar components (ρ, ρu_x, ρu_y, ρE, ϕ), each a DG Q1 function
# numbered cell by cell: U[k, c, l].  Plain loops, small meshes.
# ---------------------------------------------------------------------------
def euler_pressure(rho, rhou, rhoE, gamma):
    """pressure (covo.jl:185-190): (γ-1)(ρE - tr(ρu ⊗ vel)/2)."""
    vel = rhou / rho
    return (gamma - 1) * (rhoE - (rhou[0] * vel[0] + rhou[1] * vel[1]) / 2)


def euler_volume_flux(w, gamma):
    """The fluxes of _flux_Ω (covo.jl:44-58): rows for ρ, ρu_x, ρu_y, ρE, ϕ,
    columns x, y (ρuu + pI for the momentum rows)."""
    rho, rhou, rhoE, phi = w[0], np.array([w[1], w[2]]), w[3], w[4]
    vel = rhou / rho
    p = euler_pressure(rho, rhou, rhoE, gamma)
    F = np.zeros((5, 2))
    F[0] = rhou
    F[1:3] = np.outer(rhou, vel) + p * np.eye(2)
    F[3] = (rhoE + p) * vel
    F[4] = phi * vel
    return F


def roe_flux(ui, uj, n, gamma):
    """flux_roe (covo.jl:69-150), restated line by line: the flux (ρ, ρu_x,
    ρu_y, ρE, ϕ) through a face of normal n (side⁻'s) from state ui (side⁻)
    and uj (side⁺)."""
    nx, ny = n
    rho1, rux1, ruy1, rE1, phi1 = ui
    rho2, rux2, ruy2, rE2, phi2 = uj
    u1, v1 = rux1 / rho1, ruy1 / rho1
    u2, v2 = rux2 / rho2, ruy2 / rho2
    p1 = euler_pressure(rho1, np.array([rux1, ruy1]), rE1, gamma)
    p2 = euler_pressure(rho2, np.array([rux2, ruy2]), rE2, gamma)
    H2 = (gamma / (gamma - 1)) * p2 / rho2 + (u2 * u2 + v2 * v2) / 2.0
    H1 = (gamma / (gamma - 1)) * p1 / rho1 + (u1 * u1 + v1 * v1) / 2.0
    R = np.sqrt(rho1 / rho2)
    invR1 = 1.0 / (R + 1)
    uAv = (R * u1 + u2) * invR1
    vAv = (R * v1 + v2) * invR1
    Hav = (R * H1 + H2) * invR1
    cAv = np.sqrt(abs((gamma - 1) * (Hav - (uAv * uAv + vAv * vAv) / 2.0)))
    ecAv = (uAv * uAv + vAv * vAv) / 2.0
    l1 = nx * uAv + ny * vAv
    l3 = l1 + cAv
    l4 = l1 - cAv
    d1 = rho1 - rho2
    d2 = rho1 * u1 - rho2 * u2
    d3 = rho1 * v1 - rho2 * v2
    d4 = rE1 - rE2
    flu11 = nx * rho2 * u2 + ny * rho2 * v2
    flu21 = nx * p2 + flu11 * u2
    flu31 = ny * p2 + flu11 * v2
    flu41 = H2 * flu11
    rc1 = (gamma - 1) / cAv
    rc2 = (gamma - 1) / cAv / cAv
    uq41 = ecAv / cAv + cAv / (gamma - 1)
    uq42 = nx * uAv + ny * vAv
    fdc1 = max(l1, 0.0) * (d1 + rc2 * (-ecAv * d1 + uAv * d2 + vAv * d3 - d4))
    fdc2 = max(l1, 0.0) * ((nx * vAv - ny * uAv) * d1 + ny * d2 - nx * d3)
    fdc3 = max(l3, 0.0) * ((-uq42 * d1 + nx * d2 + ny * d3) / 2.0
                           + rc1 * (ecAv * d1 - uAv * d2 - vAv * d3 + d4) / 2.0)
    fdc4 = max(l4, 0.0) * ((uq42 * d1 - nx * d2 - ny * d3) / 2.0
                           + rc1 * (ecAv * d1 - uAv * d2 - vAv * d3 + d4) / 2.0)
    duv1 = fdc1 + (fdc3 + fdc4) / cAv
    duv2 = uAv * fdc1 + ny * fdc2 + (uAv / cAv + nx) * fdc3 + (uAv / cAv - nx) * fdc4
    duv3 = vAv * fdc1 - nx * fdc2 + (vAv / cAv + ny) * fdc3 + (vAv / cAv - ny) * fdc4
    duv4 = (ecAv * fdc1 + (ny * uAv - nx * vAv) * fdc2 + (uq41 + uq42) * fdc3
            + (uq41 - uq42) * fdc4)
    v12 = 0.5 * ((u1 + u2) * nx + (v1 + v2) * ny)
    fphi = max(0.0, v12) * phi1 + min(0.0, v12) * phi2
    return np.array([flu11 + duv1, flu21 + duv2, flu31 + duv3, flu41 + duv4, fphi])


_TRI_COORDS = np.array([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0]])              # shape.jl:207
_TRI_F2N = [(0, 1), (1, 2), (2, 0)]                                        # shape.jl:253
_TRI_NORMALS = np.array([[0.0, -1.0], [1.0, 1.0] / np.sqrt(2.0), [-1.0, 0.0]])  # shape.jl:277


def _shape2d(mesh, c):
    """(geometry name, shape name, corners, faces2nodes, reference normals) of cell c."""
    if int(mesh.celltypes[c]) == CT["Tri3"]:
        return "Tri3", "Triangle", _TRI_COORDS, _TRI_F2N, _TRI_NORMALS
    return "Quad4", "Square", _SQ_COORDS, _SQ_F2N, _SQ_NORMALS


def faces_2d(mesh: OMesh):
    """_build_faces! (mesh.jl:328-374) on a Quad4/Tri3 mesh: [cell_n, side_n,
    cell_p, side_p] in creation order, cell_p = -1 on the boundary."""
    faces, index = [], {}
    for c in range(len(mesh.celltypes)):
        nodes = mesh.c2n[mesh.c2n_ptr[c]:mesh.c2n_ptr[c + 1]]
        for s, (a, b) in enumerate(_shape2d(mesh, c)[3]):
            key = tuple(sorted((int(nodes[a]), int(nodes[b]))))
            if key in index:
                faces[index[key]][2:] = [c, s]
            else:
                index[key] = len(faces)
                faces.append([c, s, -1, -1])
    return faces


def tag_boundary_faces(mesh: OMesh, nodes_by_tag):
    """Boundary faces per tag, (cell, side) in face order: a face with one
    cell whose two nodes both lie in the tag's node set, first matching tag
    (_build_boundary_faces!, mesh.jl:383-423)."""
    out = {name: [] for name in nodes_by_tag}
    sets = {name: set(int(v) for v in nodes) for name, nodes in nodes_by_tag.items()}
    for cn, sn, cp, _ in faces_2d(mesh):
        if cp >= 0:
            continue
        nodes = mesh.c2n[mesh.c2n_ptr[cn]:mesh.c2n_ptr[cn + 1]]
        a, b = (int(nodes[k]) for k in _shape2d(mesh, cn)[3][sn])
        for name, s in sets.items():
            if a in s and b in s:
                out[name].append((cn, sn))
                break
    return out


def periodic_faces(mesh: OMesh, nodes_by_tag, shift, master, slave, tol=1e-9):
    """_compute_periodicity (domain.jl:269-397) and the periodic FaceInfo
    (domain.jl:845-890): for each slave face (in face order) the master face
    whose centre translated by `shift` is the slave face's centre; side⁻ is
    the slave cell, side⁺ the master cell.  Returns [cell⁻, side⁻, cell⁺,
    side⁺, flip] with flip = 1 when the master side's first node maps onto the
    slave side's second node (side⁺ traversed against side⁻)."""
    X = mesh.coords.reshape(-1, 2)
    tags = tag_boundary_faces(mesh, nodes_by_tag)
    f1 = [f for lab in master for f in tags[lab]]
    f2 = [f for lab in slave for f in tags[lab]]

    def face_nodes(c, s):
        nodes = mesh.c2n[mesh.c2n_ptr[c]:mesh.c2n_ptr[c + 1]]
        return [int(nodes[k]) for k in _shape2d(mesh, c)[3][s]]

    out = []
    for ci, si in f2:
        ni = face_nodes(ci, si)
        Mi = X[ni].mean(axis=0)
        cnodes = mesh.c2n[mesh.c2n_ptr[ci]:mesh.c2n_ptr[ci + 1]]
        dx = np.linalg.norm(X[cnodes].mean(axis=0) - Mi) * np.sqrt(2.0)
        for cj, sj in f1:
            nj = face_nodes(cj, sj)
            Mj = X[nj].mean(axis=0) + shift
            if np.linalg.norm(Mi - Mj) <= tol * dx:
                flip = 0 if np.linalg.norm(X[nj[0]] + shift - X[ni[0]]) <= tol * dx else 1
                out.append([ci, si, cj, sj, flip])
                break
        else:
            raise ValueError("no periodic counterpart")
    return out


def _side_eval2(mesh, c, xi_cell, side):
    """Basis values (padded to 4), unit normal of `side` and the physical point
    of cell c (Quad4 or Tri3) at reference point xi_cell."""
    geo, shp, _, _, nref = _shape2d(mesh, c)
    nodes = mesh.c2n[mesh.c2n_ptr[c]:mesh.c2n_ptr[c + 1]]
    X = mesh.coords.reshape(-1, 2)[nodes]
    J, Ji, det, x = geometry(geo, X, xi_cell)
    val, _ = shape_functions(shp, 1, xi_cell)
    n = nref[side] @ Ji
    return np.pad(val, (0, 4 - len(val))), n / np.linalg.norm(n), x


def euler_dg_residual(mesh: OMesh, U, gamma, degree, periodic=(), terms=("vol", "face")):
    """b = ∫ flux_Ω(u, v) dΩ - Σ_domains ∫ flux_Γ(u, v, n) dΓ of covo.jl:14-33
    for DG P1/Q1 on Tri3/Quad4 cells: U and the result are (5, ncells, 4)
    (the 4th entry unused on triangles).  Interior faces come from
    _build_faces!, `periodic` lists [cell⁻, side⁻, cell⁺, side⁺, flip] faces
    (periodic_faces)."""
    nc = len(mesh.celltypes)
    X = mesh.coords.reshape(-1, 2)
    b = np.zeros((5, nc, 4))
    if "vol" in terms:
        for c in range(nc):
            geo, shp = _shape2d(mesh, c)[:2]
            wv, xv = quadrature(shp, degree)
            xv = xv.reshape(-1, 2)[:len(wv)]
            nodes = mesh.c2n[mesh.c2n_ptr[c]:mesh.c2n_ptr[c + 1]]
            nl = 3 if geo == "Tri3" else 4
            for w, xi in zip(wv, xv):
                J, Ji, det, _ = geometry(geo, X[nodes], xi)
                val, gref = shape_functions(shp, 1, xi)
                grad = gref.reshape(nl, 2) @ Ji
                wq = U[:, c, :nl] @ val
                F = euler_volume_flux(wq, gamma)
                b[:, c, :nl] += w * abs(det) * (F @ grad.T)
    if "face" in terms:
        wq_, xq_ = quadrature("Line", degree)
        xq_ = xq_.reshape(-1)[:len(wq_)]
        faces = [[cn, sn, cp, sp, None] for cn, sn, cp, sp in faces_2d(mesh) if cp >= 0]
        faces += [list(f) for f in periodic]
        for cn, sn, cp, sp, flip in faces:
            nodes_n = mesh.c2n[mesh.c2n_ptr[cn]:mesh.c2n_ptr[cn + 1]]
            f2n_n, f2n_p = _shape2d(mesh, cn)[3], _shape2d(mesh, cp)[3]
            fa, fb = f2n_n[sn]
            dS = np.linalg.norm(X[nodes_n[fb]] - X[nodes_n[fa]]) / 2
            if flip is None:   # interior face: indexin(f2n_p, f2n_n)
                nodes_p = mesh.c2n[mesh.c2n_ptr[cp]:mesh.c2n_ptr[cp + 1]]
                gn = [int(nodes_n[k]) for k in f2n_n[sn]]
                gp = [int(nodes_p[k]) for k in f2n_p[sp]]
                perm = [gn.index(v) for v in gp]
                loc_p = tuple(f2n_p[sp][k] for k in perm)
            else:
                a, bb = f2n_p[sp]
                loc_p = (bb, a) if flip else (a, bb)
            Vn, Vp = _shape2d(mesh, cn)[2], _shape2d(mesh, cp)[2]
            for w, t in zip(wq_, xq_):
                xn = 0.5 * (1.0 - t) * Vn[f2n_n[sn][0]] + 0.5 * (1.0 + t) * Vn[f2n_n[sn][1]]
                xp = 0.5 * (1.0 - t) * Vp[loc_p[0]] + 0.5 * (1.0 + t) * Vp[loc_p[1]]
                vn, nn, _ = _side_eval2(mesh, cn, xn, sn)
                vp, _, _ = _side_eval2(mesh, cp, xp, sp)
                F = roe_flux(U[:, cn, :] @ vn, U[:, cp, :] @ vp, nn, gamma)
                b[:, cn, :] -= w * dS * np.outer(F, vn)
                b[:, cp, :] += w * dS * np.outer(F, vp)
    return b


def euler_dg_mass_inverse_apply(mesh: OMesh, R, degree):
    """M⁻¹ R with M the DG P1/Q1 mass matrix (block-diagonal, one block per
    cell and component): the `cache.mass \\ rhs` of covo.jl:31."""
    nc = len(mesh.celltypes)
    X = mesh.coords.reshape(-1, 2)
    out = np.zeros_like(R)
    for c in range(nc):
        geo, shp = _shape2d(mesh, c)[:2]
        nl = 3 if geo == "Tri3" else 4
        wv, xv = quadrature(shp, degree)
        xv = xv.reshape(-1, 2)[:len(wv)]
        nodes = mesh.c2n[mesh.c2n_ptr[c]:mesh.c2n_ptr[c + 1]]
        M = np.zeros((nl, nl))
        for w, xi in zip(wv, xv):
            _, _, det, _ = geometry(geo, X[nodes], xi)
            val, _ = shape_functions(shp, 1, xi)
            M += w * abs(det) * np.outer(val, val)
        out[:, c, :nl] = np.linalg.solve(M, R[:, c, :nl].T).T
    return out
