"""Function spaces and FE spaces on top of libbcubehip.

Mirrors FunctionSpace (src/function_space/function_space.jl, Lagrange
lagrange.jl:16), DofHandler (src/fespace/dofhandler.jl:37-156) and
TrialFESpace / TestFESpace (src/fespace/fespace.jl:186-414).  The dof numbering
is computed by the library (bh_dofhandler), bit-identical to the reference's.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .mesh import Mesh


class FunctionSpace:
    """FunctionSpace(:Lagrange, degree) -- uniform-node Lagrange (Lagrange{:Uniform}).

    `degree` may also be one order (1..4) per cell of a 2D Quad4/Tri3 mesh:
    the continuous space of per-cell order whose edges take their cells'
    lowest order (bh_space_create_orders; no counterpart in the reference,
    whose FunctionSpace has one degree).  `degree` is then the highest order
    and `cell_degrees` the array."""

    def __init__(self, kind: str, degree):
        if kind.lower() != "lagrange":
            raise NotImplementedError(f"function space {kind!r}: only Lagrange is supported")
        self.kind = "Lagrange"
        self.cell_degrees = None
        if np.ndim(degree) > 0:
            d = np.ascontiguousarray(degree, dtype=np.int32).reshape(-1)
            if d.size == 0 or d.min() < 1 or d.max() > 4:
                raise ValueError("per-cell degrees must be 1..4")
            self.cell_degrees = d
            self.degree = int(d.max())
            return
        if degree < 0:
            raise ValueError("degree must be >= 0")
        self.degree = int(degree)

    def __repr__(self):
        if self.cell_degrees is not None:
            return f"FunctionSpace(:{self.kind}, per-cell degrees {np.bincount(self.cell_degrees)[1:]})"
        return f"FunctionSpace(:{self.kind}, {self.degree})"


def component_numbering(iglob, offset, ncomp: int) -> np.ndarray:
    """Global ids, 0-based, of component k's copy of scalar dof g, as an
    (ncomp, ndofs) array: the numbering DofHandler(mesh, fs, ncomp, ...) gives
    (dofhandler.jl:37-156).  There the cells' (component, local dof) slots are
    numbered in order of first appearance, the blocks of a cell's components
    following each other, and a component's dof on a shared vertex, edge or
    face copies the first cell's id of the same component.  So the dofs that
    first appear in cell c -- n_c of them, following the base(c) dofs of the
    cells before, with scalar ids base(c) + r in local order -- take
    ncomp*base(c) + k*n_c + r."""
    iglob = np.asarray(iglob, dtype=np.int64) - 1
    n = int(iglob.max()) + 1 if iglob.size else 0
    first = np.full(n, iglob.size, dtype=np.int64)
    np.minimum.at(first, iglob, np.arange(iglob.size, dtype=np.int64))
    cell = np.searchsorted(np.asarray(offset, dtype=np.int64), first, side="right") - 1
    ncell = np.bincount(cell, minlength=len(offset))
    base = np.concatenate([[0], np.cumsum(ncell)])[:-1]
    g = np.arange(n, dtype=np.int64)
    r = g - base[cell]
    k = np.arange(ncomp, dtype=np.int64)[:, None]
    return ncomp * base[cell][None, :] + k * ncell[cell][None, :] + r[None, :]


class DofHandler:
    """Global dof numbering of a Lagrange space of `ncomponents` components.

    iglob: 1-based global ids, concatenated per cell (the components' blocks
    following each other inside a cell); offset[c] locates cell c, and
    offset[c] + k * ndofs(c) its component k.
    """

    def __init__(self, mesh: Mesh, fs: FunctionSpace, ncomponents: int = 1,
                 isContinuous: bool = True, device: bool = False):
        """device=True numbers on the GPU (bh_dofhandler_device, the same
        ids), for meshes that change; it needs a mesh on a HIP backend."""
        if ncomponents < 1:
            raise ValueError("ncomponents must be >= 1")
        self.ncomponents = int(ncomponents)
        self.mesh = mesh
        self.fs = fs
        self.continuous = bool(isContinuous)
        nd = {0: (fs.degree + 1) ** 2, 1: (fs.degree + 1) ** 2,
              2: (fs.degree + 1) * (fs.degree + 2) // 2, 3: (fs.degree + 1) ** 3,
              4: (fs.degree + 1) ** 3, 5: fs.degree + 1}
        total = int(sum(nd[int(t)] for t in mesh.celltypes))
        self.iglob = np.zeros(total, dtype=np.int64)
        self.offset = np.zeros(mesh.ncells, dtype=np.int64)
        n = C.c_int64()
        if device:
            ctx = mesh.backend.ctx
            rc = L.lib().bh_dofhandler_device(mesh.handle, fs.degree, int(self.continuous),
                                              self.iglob.ctypes.data, self.offset.ctypes.data,
                                              C.byref(n))
            L.check(rc, ctx, "bh_dofhandler_device")
        else:
            rc = L.lib().bh_dofhandler(_host_mesh(mesh), fs.degree, int(self.continuous),
                                       self.iglob.ctypes.data, self.offset.ctypes.data,
                                       C.byref(n))
            L.check(rc, None, "bh_dofhandler")  # host-only handle: no GPU context needed
        self.ndofs_tot = int(n.value)
        self._ndofs_cell = np.diff(np.append(self.offset, total))
        if self.ncomponents > 1:
            # the scalar numbering above, expanded to the components
            nc = self.ncomponents
            self.scalar_iglob, self.scalar_offset = self.iglob, self.offset
            self.component_map = component_numbering(self.iglob, self.offset, nc)
            cell = np.repeat(np.arange(mesh.ncells), self._ndofs_cell)
            loc = np.arange(total) - self.offset[cell]
            vec = self.component_map[:, self.iglob - 1] + 1  # (nc, total)
            self.offset = self.offset * nc
            pos = self.offset[cell][None, :] + np.arange(nc)[:, None] * self._ndofs_cell[cell] + loc
            self.iglob = np.empty(total * nc, dtype=np.int64)
            self.iglob[pos.reshape(-1)] = vec.reshape(-1)
            self.ndofs_tot *= nc

    def get_ndofs(self, icell: int | None = None) -> int:
        if icell is None:
            return self.ndofs_tot
        return int(self._ndofs_cell[icell]) * self.ncomponents

    def get_dof(self, icell: int, icomp: int | None = None) -> np.ndarray:
        """1-based global dofs of cell icell (0-based cell index), all
        components or component icomp (0-based) (get_dof, dofhandler.jl:489-505)."""
        o = self.offset[icell]
        nd = self._ndofs_cell[icell]
        if icomp is None:
            return self.iglob[o:o + nd * self.ncomponents]
        return self.iglob[o + icomp * nd:o + (icomp + 1) * nd]


def _host_mesh(mesh: Mesh):
    """A context-free library mesh for numbering queries (no GPU needed),
    cached on the mesh and freed with it."""
    h = getattr(mesh, "_host_handle", None)
    if h is None:
        h = C.c_void_p()
        L.check(L.lib().bh_mesh_create(None, mesh.dim, mesh.nnodes, mesh.coords.ctypes.data,
                                       mesh.ncells, mesh.celltypes.ctypes.data,
                                       mesh.c2n_ptr.ctypes.data, mesh.c2n.ctypes.data,
                                       C.byref(h)), None, "bh_mesh_create")
        mesh._host_handle = h
    return h


class SingleFESpace:
    """A Lagrange FE space of `size` components with a device-resident per-cell
    dof map.  The library works on the scalar numbering (`handle`); a vector
    space (size > 1) adds the map of each component's copy of the scalar dofs
    into its own numbering (`component_map`)."""

    def __init__(self, fs: FunctionSpace, mesh: Mesh, isContinuous: bool = True, size: int = 1):
        self.fs = fs
        self.mesh = mesh
        self.continuous = bool(isContinuous)
        self.size = int(size)
        self._handle = None
        if fs.cell_degrees is not None:
            # continuous space of per-cell order: numbered by the library
            if not self.continuous or self.size != 1:
                raise NotImplementedError("per-cell degrees: continuous scalar spaces only")
            if fs.cell_degrees.size != mesh.ncells:
                raise ValueError("one degree per cell")
            h = C.c_void_p()
            L.check(L.lib().bh_space_create_orders(mesh.handle, fs.cell_degrees.ctypes.data,
                                                   C.byref(h)), mesh.backend.ctx,
                    "bh_space_create_orders")
            self._handle = h
            n, nl = C.c_int64(), C.c_int32()
            L.check(L.lib().bh_space_info(h, C.byref(n), C.byref(nl)), mesh.backend.ctx,
                    "bh_space_info")
            self._ndofs, self._nl = n.value, nl.value
            self.dhl = self._scalar = None
            return
        self.dhl = DofHandler(mesh, fs, self.size, self.continuous)
        self._scalar = DofHandler(mesh, fs, 1, self.continuous) if self.size > 1 else self.dhl

    def get_ndofs(self) -> int:
        if self.dhl is None:
            return self._ndofs
        return self.dhl.ndofs_tot

    def get_scalar_ndofs(self) -> int:
        if self._scalar is None:
            return self._ndofs
        return self._scalar.ndofs_tot

    def cell_slots(self) -> np.ndarray:
        """(ncells, nl) 0-based dof of each cell's Lagrange slot, -1 where the
        slot holds no dof (the library's per-cell dof map)."""
        out = np.empty(self.mesh.ncells * self.nl_slots, dtype=np.int32)
        L.check(L.lib().bh_space_dofs(self.handle, out.ctypes.data), self.mesh.backend.ctx,
                "bh_space_dofs")
        return out.reshape(self.mesh.ncells, -1)

    @property
    def nl_slots(self) -> int:
        n, nl = C.c_int64(), C.c_int32()
        L.check(L.lib().bh_space_info(self.handle, C.byref(n), C.byref(nl)),
                self.mesh.backend.ctx, "bh_space_info")
        return nl.value

    def edge_orders(self):
        """Space of per-cell order: (ncells, 4) order of each local edge of
        lower order than its cell (0 where the edge has the cell's order), and
        the number of cells with such an edge."""
        eo = np.zeros(self.mesh.ncells * 4, dtype=np.uint8)
        nt = C.c_int64()
        L.check(L.lib().bh_space_edge_orders(self.handle, eo.ctypes.data, C.byref(nt)),
                self.mesh.backend.ctx, "bh_space_edge_orders")
        return eo.reshape(-1, 4), nt.value

    def get_dofs(self, icell: int) -> np.ndarray:
        if self.dhl is None:
            d = self.cell_slots()[icell]
            return d[d >= 0].astype(np.int64) + 1
        return self.dhl.get_dof(icell)

    @property
    def component_map(self) -> np.ndarray:
        """(size, scalar ndofs) 0-based ids of each component's dofs."""
        if self.size == 1:
            return np.arange(self._scalar.ndofs_tot, dtype=np.int64)[None, :]
        return self.dhl.component_map

    @property
    def handle(self):
        """bh_space on the mesh's backend, numbered from the scalar DofHandler."""
        if self._handle is None:
            h = C.c_void_p()
            m = self.mesh.handle
            sd = self._scalar
            L.check(L.lib().bh_space_create_from_dofs(
                m, self.fs.degree, int(self.continuous), sd.ndofs_tot,
                sd.iglob.ctypes.data, sd.offset.ctypes.data, C.byref(h)),
                self.mesh.backend.ctx, "bh_space_create_from_dofs")
            self._handle = h
        return self._handle

    def __del__(self):
        try:
            if self._handle is not None:
                L.lib().bh_space_destroy(self._handle)
        except Exception:
            pass


class TrialFESpace(SingleFESpace):
    """TrialFESpace(fs, mesh[, :discontinuous][, dirichlet]) (fespace.jl:186-274):
    `dirichlet` maps boundary tags to values (a number, a PhysicalFunction, an
    FEFunction or a function of time), applied by the functions of
    bcubehip.dirichlet."""

    def __init__(self, fs: FunctionSpace, mesh: Mesh, kind="continuous", dirichlet=None, *,
                 isContinuous: bool | None = None, size: int = 1):
        if isinstance(kind, dict):  # TrialFESpace(fs, mesh, Dict(...))
            kind, dirichlet = "continuous", kind
        cont = (kind != "discontinuous") if isContinuous is None else bool(isContinuous)
        super().__init__(fs, mesh, cont, size)
        self.dirichlet = dict(dirichlet or {})
        if self.dirichlet and self.size > 1:
            raise NotImplementedError("Dirichlet conditions on vector spaces")


class TestFESpace:
    """TestFESpace(U) (fespace.jl:414): shares U's numbering.  Also
    TestFESpace(fs, mesh[, kind]; size), a space of its own."""

    __test__ = False  # not a pytest class

    def __init__(self, U, mesh: Mesh | None = None, kind="continuous", *, size: int = 1):
        if isinstance(U, FunctionSpace):
            U = TrialFESpace(U, mesh, kind, size=size)
        self.U = U

    def __getattr__(self, name):
        return getattr(self.U, name)


def _single(space):
    return getattr(space, "U", space)


class MultiFESpace:
    """MultiFESpace(U1, U2, ...; arrayOfStruct=True) (fespace.jl:477-647): one
    global numbering over several FE spaces.  get_mapping(i)[g] is the global
    id (1-based) of space i's dof g (0-based), numbered by first appearance
    over cells then spaces (AoS, _build_mapping_AoS) or spaces then cells
    (SoA, _build_mapping_SoA)."""

    __test__ = False

    def __init__(self, *spaces, arrayOfStruct: bool = True):
        if len(spaces) == 1 and isinstance(spaces[0], (tuple, list)):
            spaces = tuple(spaces[0])
        self.spaces = tuple(spaces)
        self.arrayOfStruct = bool(arrayOfStruct)
        singles = [_single(s) for s in self.spaces]
        ncells = singles[0].mesh.ncells
        if any(s.mesh.ncells != ncells for s in singles):
            raise ValueError("the spaces of a MultiFESpace share one mesh")
        # every (cell, space, local dof) slot in traversal order; a dof takes
        # the rank of its first slot
        keyoff = np.concatenate([[0], np.cumsum([s.get_ndofs() for s in singles])])
        cols = {"cell": [], "space": [], "loc": [], "key": []}
        for i, sp in enumerate(singles):
            dh = sp.dhl
            nd = dh._ndofs_cell * dh.ncomponents
            cell = np.repeat(np.arange(ncells), nd)
            cols["cell"].append(cell)
            cols["space"].append(np.full(cell.size, i))
            cols["loc"].append(np.arange(cell.size) - np.repeat(dh.offset, nd))
            cols["key"].append(keyoff[i] + dh.iglob - 1)
        cell, space, loc, key = (np.concatenate(cols[k]) for k in ("cell", "space", "loc", "key"))
        order = (np.lexsort((loc, space, cell)) if self.arrayOfStruct
                 else np.lexsort((loc, cell, space)))
        uniq, first = np.unique(key[order], return_index=True)
        ids = np.zeros(int(keyoff[-1]), dtype=np.int64)
        ids[uniq[np.argsort(first)]] = np.arange(1, uniq.size + 1)
        if uniq.size != keyoff[-1]:
            raise ValueError("invalid mapping")
        self.mapping = [ids[keyoff[i]:keyoff[i + 1]] for i in range(len(singles))]
        self.ndofs = int(uniq.size)

    def __iter__(self):
        return iter(self.spaces)

    def __len__(self):
        return len(self.spaces)

    def __getitem__(self, i):
        return self.spaces[i]

    def get_ndofs(self) -> int:
        return self.ndofs

    def get_mapping(self, i: int) -> np.ndarray:
        """get_mapping(U, i) (fespace.jl:610-647), for 0-based i."""
        return self.mapping[i]

    def get_fespace(self):
        return self.spaces
