"""Dirichlet conditions on tagged boundaries (src/assembler/dirichlet_condition.jl).

As in the reference, the host side walks the trial space's Dirichlet
conditions: for every boundary tag it takes the tag's boundary faces
(`BoundaryFaceDomain`, faces tagged by `_build_boundary_faces!`,
mesh.jl:383-423), the dofs of each face's cell lying on that face, vertices
and edges included (`idof_by_face_with_bounds`), and the condition's value at
each dof (`_apply_dirichlet_on_face!`, dirichlet_condition.jl:264-330).  The
device side is the C ABI (`bh_dirichlet_*`): the CSC value positions of those
dofs' rows, columns and diagonal, computed once, and the edits applied in
place on host or device arrays.

Values: a number, a `PhysicalFunction` (evaluated at the dof's physical
point), an `FEFunction` of a continuous space of the same mesh and order
(its dof values), or a function of time returning one of these.  Tags apply in
the order of the space's Dirichlet dict; a dof shared by two tags takes the
later tag's value.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .forms import FEFunction, PhysicalFunction
from .mesh import _H8_VERTS, _H27_REF

_QUAD4, _TRI3, _HEXA8 = L.CELLTYPE["Quad4"], L.CELLTYPE["Tri3"], L.CELLTYPE["Hexa8"]
_QUAD9, _HEXA27 = L.CELLTYPE["Quad9"], L.CELLTYPE["Hexa27"]

# sides of each cell type as local vertex lists, and the reference vertices
_SIDES = {
    _QUAD4: [(0, 1), (1, 2), (2, 3), (3, 0)],
    _TRI3: [(0, 1), (1, 2), (2, 0)],
    _HEXA8: [(0, 3, 2, 1), (0, 1, 5, 4), (1, 2, 6, 5), (2, 3, 7, 6), (0, 4, 7, 3), (4, 5, 6, 7)],
}
_VREF = {
    _QUAD4: np.array([[-1, -1], [1, -1], [1, 1], [-1, 1]], dtype=np.float64),
    _TRI3: np.array([[0, 0], [1, 0], [0, 1]], dtype=np.float64),
    _HEXA8: _H8_VERTS.astype(np.float64),
}
# curved cells: sides and vertices are those of the straight cell of the same
# shape (the first nodes in CGNS order); the geometry uses every node
_SIDES[_QUAD9] = _SIDES[_QUAD4]
_SIDES[_HEXA27] = _SIDES[_HEXA8]
_VREF[_QUAD9] = _VREF[_QUAD4]
_VREF[_HEXA27] = _VREF[_HEXA8]
# reference coordinates of every geometric node of the curved cells, CGNS order
# (entity.jl: Quad9 vertices, edge midpoints 12 23 34 41, centre; Hexa27
# vertices, 12 edges, 6 face centres, centre)
_GREF = {
    _QUAD9: np.array([[-1, -1], [1, -1], [1, 1], [-1, 1], [0, -1], [1, 0], [0, 1], [-1, 0],
                      [0, 0]], dtype=np.float64),
}
_GREF[_HEXA27] = np.asarray(_H27_REF, dtype=np.float64)
# triangle Lagrange nodes as barycentric lattice points (i, j, k), i + j + k = p,
# in the reference's dof order (triangle.jl:1-45, 120-145; P4 as the library)
_TRI_LATTICE = {
    1: [(1, 0, 0), (0, 1, 0), (0, 0, 1)],
    2: [(2, 0, 0), (0, 2, 0), (0, 0, 2), (1, 1, 0), (0, 1, 1), (1, 0, 1)],
    3: [(3, 0, 0), (0, 3, 0), (0, 0, 3), (2, 1, 0), (1, 2, 0), (0, 2, 1), (0, 1, 2), (1, 0, 2),
        (2, 0, 1), (1, 1, 1)],
    4: [(4, 0, 0), (0, 4, 0), (0, 0, 4), (3, 1, 0), (2, 2, 0), (1, 3, 0), (0, 3, 1), (0, 2, 2),
        (0, 1, 3), (1, 0, 3), (2, 0, 2), (3, 0, 1), (2, 1, 1), (1, 2, 1), (1, 1, 2)],
}


def reference_nodes(celltype: int, p: int) -> np.ndarray:
    """Reference coordinates of the Lagrange dofs of one cell, in dof order
    (tensor cells: uniform nodes on [-1, 1], first index fastest,
    hypercube.jl:63, 85; triangles: the lattice above on the unit triangle)."""
    if celltype == _TRI3:
        lat = np.array(_TRI_LATTICE[p], dtype=np.float64)
        return lat[:, 1:] / p
    x = -1.0 + 2.0 * np.arange(p + 1) / p
    if celltype in (_QUAD4, _QUAD9):
        return np.array([[x[i], x[j]] for j in range(p + 1) for i in range(p + 1)])
    if celltype in (_HEXA8, _HEXA27):
        return np.array([[x[i], x[j], x[k]] for k in range(p + 1) for j in range(p + 1)
                         for i in range(p + 1)])
    raise NotImplementedError("Dirichlet conditions: Quad4, Quad9, Tri3, Hexa8 and Hexa27 cells")


def _quadratic_1d(s: np.ndarray, t: np.ndarray) -> np.ndarray:
    """The 1D quadratic Lagrange basis on {-1, 0, 1}: the function of the node s at t."""
    return np.where(s == -1, 0.5 * t * (t - 1.0), np.where(s == 0, 1.0 - t * t,
                                                          0.5 * t * (t + 1.0)))


def _geometric_weights(celltype: int, xi: np.ndarray) -> np.ndarray:
    """(npts, ngeometric nodes) values of the geometric basis: affine /
    multilinear on straight cells, products of 1D quadratics on Quad9 / Hexa27
    (mapping.jl:544-570, 664-697)."""
    if celltype == _TRI3:
        return np.stack([1.0 - xi[:, 0] - xi[:, 1], xi[:, 0], xi[:, 1]], axis=1)
    if celltype in _GREF:
        g = _GREF[celltype]
        return np.prod(_quadratic_1d(g[None, :, :], xi[:, None, :]), axis=2)
    v = _VREF[celltype]
    return np.prod((1.0 + xi[:, None, :] * v[None, :, :]) / 2.0, axis=2)


def _side_nodes(celltype: int, side: int, xi: np.ndarray) -> np.ndarray:
    """Local dofs (reference points xi) lying on a side, bounds included."""
    verts = list(_SIDES[celltype][side])
    if celltype == _TRI3:
        c = ({0, 1, 2} - set(verts)).pop()  # the side has no weight on its opposite vertex
        lam = np.stack([1.0 - xi[:, 0] - xi[:, 1], xi[:, 0], xi[:, 1]], axis=1)
        return np.flatnonzero(np.isclose(lam[:, c], 0.0))
    v = _VREF[celltype][verts]
    a = int(np.flatnonzero(np.all(v == v[0], axis=0))[0])  # the side's constant axis
    return np.flatnonzero(np.isclose(xi[:, a], v[0, a]))


def _cell_vertices(mesh, icell: int) -> np.ndarray:
    nodes = mesh.cell_nodes(icell)
    return nodes[:len(_VREF[int(mesh.celltypes[icell])])]


def _geometric_nodes(mesh, icell: int) -> np.ndarray:
    """The nodes the geometric map of a cell uses: all of a curved cell's."""
    ct = int(mesh.celltypes[icell])
    return mesh.cell_nodes(icell) if ct in _GREF else _cell_vertices(mesh, icell)


def boundary_faces(mesh) -> dict[str, list[tuple[int, int]]]:
    """Boundary faces of each tag as (cell, side), in face order (the creating
    cell, then its local side: `_build_faces!`, mesh.jl:328-374).  A face with
    one cell takes the first tag whose node set holds all its nodes
    (`_build_boundary_faces!`, mesh.jl:383-423)."""
    count: dict[tuple, int] = {}
    sides = []
    for c in range(mesh.ncells):
        ct = int(mesh.celltypes[c])
        if ct not in _SIDES:
            raise NotImplementedError("Dirichlet conditions: Quad4, Quad9, Tri3, Hexa8 and "
                                      "Hexa27 cells")
        vn = _cell_vertices(mesh, c)
        for s, loc in enumerate(_SIDES[ct]):
            key = tuple(sorted(int(vn[k]) for k in loc))
            count[key] = count.get(key, 0) + 1
            sides.append((c, s, key))
    member = {}
    for name, nodes in mesh.bc_nodes.items():
        mask = np.zeros(mesh.nnodes, dtype=bool)
        mask[np.asarray(nodes, dtype=np.int64)] = True
        member[name] = mask
    out: dict[str, list[tuple[int, int]]] = {name: [] for name in mesh.bc_nodes}
    for c, s, key in sides:
        if count[key] != 1:
            continue
        for name, mask in member.items():
            if mask[list(key)].all():
                out[name].append((c, s))
                break
    return out


def dof_coordinates(U) -> np.ndarray:
    """Physical points of the dofs of a continuous Lagrange space, (ndofs, dim):
    the cell map of each dof's reference node (all cells sharing a dof give
    the same point)."""
    U = getattr(U, "U", U)
    mesh, p = U.mesh, U.fs.degree
    X = mesh.nodes()
    pts = np.zeros((U.get_ndofs(), mesh.dim))
    for ct in np.unique(mesh.celltypes):
        ct = int(ct)
        xi = reference_nodes(ct, p)
        w = _geometric_weights(ct, xi)
        cells = np.flatnonzero(mesh.celltypes == ct)
        vert = np.stack([_geometric_nodes(mesh, c) for c in cells])        # (n, nv)
        x = np.einsum("qk,ckd->cqd", w, X[vert])                            # (n, nl, dim)
        off = U.dhl.offset[cells][:, None] + np.arange(xi.shape[0])[None, :]
        pts[U.dhl.iglob[off] - 1] = x
    return pts


def _evaluate(value, t, U, dofs, pts) -> np.ndarray:
    if isinstance(value, (int, float, np.floating, np.integer)):
        return np.full(dofs.size, float(value))
    if isinstance(value, PhysicalFunction):
        return np.asarray(value.f(pts[dofs]), dtype=np.float64).reshape(-1)
    if isinstance(value, FEFunction):
        S = value.space
        if S.mesh is not U.mesh or S.fs.degree != U.fs.degree or not S.continuous:
            raise NotImplementedError("FEFunction Dirichlet values: a continuous space of the "
                                      "same mesh and order")
        return np.asarray(value.values, dtype=np.float64).reshape(-1)[dofs]
    if callable(value):
        return _evaluate(value(t), t, U, dofs, pts)
    raise TypeError(f"unsupported Dirichlet value {type(value).__name__}")


def dirichlet_dofs_values(U, t: float = 0.0) -> tuple[np.ndarray, np.ndarray]:
    """0-based dofs and values of the space's Dirichlet conditions at time t,
    in application order (tags in dict order, faces in face order, the face's
    dofs in cell order); a dof may appear several times, the last value wins."""
    U = getattr(U, "U", U)
    cond = getattr(U, "dirichlet", None) or {}
    if not cond:
        return np.zeros(0, dtype=np.int64), np.zeros(0)
    mesh, p = U.mesh, U.fs.degree
    faces = boundary_faces(mesh)
    pts = dof_coordinates(U) if any(not isinstance(v, (int, float)) for v in cond.values()) else None
    xi = {int(ct): reference_nodes(int(ct), p) for ct in np.unique(mesh.celltypes)}
    all_d, all_v = [], []
    for tag, value in cond.items():
        if tag not in faces:
            raise KeyError(f"no boundary tag {tag!r} on this mesh")
        dl = []
        for c, s in faces[tag]:
            ct = int(mesh.celltypes[c])
            loc = _side_nodes(ct, s, xi[ct])
            dl.append(U.dhl.iglob[U.dhl.offset[c] + loc] - 1)
        d = np.concatenate(dl) if dl else np.zeros(0, dtype=np.int64)
        all_d.append(d)
        all_v.append(_evaluate(value, t, U, d, pts))
    return (np.concatenate(all_d).astype(np.int64), np.concatenate(all_v).astype(np.float64))


class DirichletPlan:
    """bh_dirichlet of a space's Dirichlet dofs on its (U, U) pattern: the CSC
    positions computed once, edits applied on host or device arrays."""

    def __init__(self, U, t: float = 0.0):
        from .assembler import pattern_of
        self.U = getattr(U, "U", U)
        self.pattern = pattern_of(self.U)
        self.dofs_list, self.values = dirichlet_dofs_values(self.U, t)
        h = C.c_void_p()
        L.check(L.lib().bh_dirichlet_create(self.pattern.handle, self.dofs_list.size,
                                            self.dofs_list.ctypes.data, C.byref(h)),
                self.pattern.ctx, "bh_dirichlet_create")
        self.handle = h
        nd, npos = C.c_int64(), C.c_int64()
        L.check(L.lib().bh_dirichlet_info(h, C.byref(nd), C.byref(npos)), self.pattern.ctx,
                "bh_dirichlet_info")
        self.ndofs, self.npos = int(nd.value), int(npos.value)

    def positions(self):
        """(dofs, src, row_pos, col_pos, diag_pos) as numpy arrays."""
        dofs = np.zeros(self.ndofs, dtype=np.int64)
        src = np.zeros(self.ndofs, dtype=np.int32)
        rows = np.zeros(self.npos, dtype=np.int64)
        cols = np.zeros(self.npos, dtype=np.int64)
        diag = np.zeros(self.ndofs, dtype=np.int64)
        L.check(L.lib().bh_dirichlet_positions(self.handle, dofs.ctypes.data, src.ctypes.data,
                                               rows.ctypes.data, cols.ctypes.data,
                                               diag.ctypes.data), self.pattern.ctx,
                "bh_dirichlet_positions")
        return dofs, src, rows, cols, diag

    def apply_matrix(self, nzval, diag: float = 1.0, homogeneous: bool = False,
                     where: int = L.HOST):
        """nzval: the CSC values of the pattern (numpy array, or a device pointer
        with where=DEVICE)."""
        ptr = nzval.ctypes.data if isinstance(nzval, np.ndarray) else nzval
        L.check(L.lib().bh_dirichlet_apply_matrix(self.handle, float(diag), int(homogeneous), ptr,
                                                  where), self.pattern.ctx,
                "bh_dirichlet_apply_matrix")

    def apply_vector(self, b, values=None, where: int = L.HOST):
        """b[dof] = value (values: None for the plan's own, or an array like the
        plan's dof list; homogeneous=zeros via values=0)."""
        bptr = b.ctypes.data if isinstance(b, np.ndarray) else b
        if isinstance(values, (int, float)) and values == 0:
            vptr = None
        else:
            v = self.values if values is None else values
            vptr = v.ctypes.data if isinstance(v, np.ndarray) else v
        L.check(L.lib().bh_dirichlet_apply_vector(self.handle, vptr, bptr, where),
                self.pattern.ctx, "bh_dirichlet_apply_vector")

    def __del__(self):
        try:
            L.lib().bh_dirichlet_destroy(self.handle)
        except Exception:
            pass


def _plan(U, t=0.0) -> DirichletPlan:
    U = getattr(U, "U", U)
    cache = U.__dict__.setdefault("_dirichlet_plans", {})
    if t not in cache:
        cache[t] = DirichletPlan(U, t)
    return cache[t]


def _matrices(matrices):
    return matrices if isinstance(matrices, tuple) else (matrices,)


def apply_dirichlet_to_matrix_(matrices, U, V, mesh, diag_values=None) -> None:
    """apply_dirichlet_to_matrix!(matrix or (matrices...), U, V, mesh; diag_value[s])
    (dirichlet_condition.jl:147-197): each Dirichlet row zeroed, diagonal set,
    columns kept.  The matrices are the CSC results of assemble_bilinear."""
    mats = _matrices(matrices)
    dv = diag_values if diag_values is not None else (1.0,) * len(mats)
    if not isinstance(dv, tuple):
        dv = (dv,)
    d = _plan(U)
    for A, v in zip(mats, dv):
        d.apply_matrix(A.data, v, homogeneous=False)


def apply_homogeneous_dirichlet_to_matrix_(matrices, U, V, mesh, diag_values=None) -> None:
    """apply_homogeneous_dirichlet_to_matrix! (dirichlet_condition.jl:42-89):
    Dirichlet rows and columns zeroed, diagonal set."""
    mats = _matrices(matrices)
    dv = diag_values if diag_values is not None else (1.0,) * len(mats)
    if not isinstance(dv, tuple):
        dv = (dv,)
    d = _plan(U)
    for A, v in zip(mats, dv):
        d.apply_matrix(A.data, v, homogeneous=True)


def apply_dirichlet_to_vector_(b: np.ndarray, U, V, mesh, t: float = 0.0) -> None:
    """apply_dirichlet_to_vector!(b, U, V, mesh, t) (dirichlet_condition.jl:118-145)."""
    _plan(U, t).apply_vector(b)


def apply_homogeneous_dirichlet_to_vector_(b: np.ndarray, U, V, mesh) -> None:
    """apply_homogeneous_dirichlet_to_vector! (dirichlet_condition.jl:24-40)."""
    _plan(U).apply_vector(b, 0)


def apply_homogeneous_dirichlet_(A, b: np.ndarray, U, V, mesh) -> None:
    """apply_homogeneous_dirichlet!(A, b, U, V, mesh) (dirichlet_condition.jl:91-116)."""
    apply_homogeneous_dirichlet_to_matrix_(A, U, V, mesh)
    apply_homogeneous_dirichlet_to_vector_(b, U, V, mesh)


def assemble_dirichlet_vector(U, V, mesh, t: float = 0.0) -> np.ndarray:
    """assemble_dirichlet_vector(U, V, mesh, t) (dirichlet_condition.jl:1-27):
    zeros except the Dirichlet dofs, which take their values."""
    U = getattr(U, "U", U)
    b = np.zeros(U.get_ndofs())
    apply_dirichlet_to_vector_(b, U, V, mesh, t)
    return b
