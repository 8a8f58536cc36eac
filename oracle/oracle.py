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
        for n in ("or_result_ndofs", "or_result_nnz"):
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
