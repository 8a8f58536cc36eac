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
    """FunctionSpace(:Lagrange, degree) -- uniform-node Lagrange (Lagrange{:Uniform})."""

    def __init__(self, kind: str, degree: int):
        if kind.lower() != "lagrange":
            raise NotImplementedError(f"function space {kind!r}: only Lagrange is supported")
        if degree < 0:
            raise ValueError("degree must be >= 0")
        self.kind = "Lagrange"
        self.degree = int(degree)

    def __repr__(self):
        return f"FunctionSpace(:{self.kind}, {self.degree})"


class DofHandler:
    """Global dof numbering of a scalar Lagrange space on a mesh.

    iglob: 1-based global ids, concatenated per cell; offset[c] locates cell c.
    """

    def __init__(self, mesh: Mesh, fs: FunctionSpace, ncomponents: int = 1,
                 isContinuous: bool = True):
        if ncomponents != 1:
            raise NotImplementedError("vector-valued spaces (size > 1) are not supported yet")
        self.mesh = mesh
        self.fs = fs
        self.continuous = bool(isContinuous)
        nd = {0: (fs.degree + 1) ** 2, 1: (fs.degree + 1) ** 2,
              2: (fs.degree + 1) * (fs.degree + 2) // 2, 3: (fs.degree + 1) ** 3,
              4: (fs.degree + 1) ** 3}
        total = int(sum(nd[int(t)] for t in mesh.celltypes))
        self.iglob = np.zeros(total, dtype=np.int64)
        self.offset = np.zeros(mesh.ncells, dtype=np.int64)
        n = C.c_int64()
        rc = L.lib().bh_dofhandler(_host_mesh(mesh), fs.degree, int(self.continuous),
                                   self.iglob.ctypes.data, self.offset.ctypes.data, C.byref(n))
        L.check(rc, None, "bh_dofhandler")  # host-only handle: no GPU context needed
        self.ndofs_tot = int(n.value)
        self._ndofs_cell = np.diff(np.append(self.offset, total))

    def get_ndofs(self, icell: int | None = None) -> int:
        if icell is None:
            return self.ndofs_tot
        return int(self._ndofs_cell[icell])

    def get_dof(self, icell: int) -> np.ndarray:
        """1-based global dofs of cell icell (0-based cell index)."""
        o = self.offset[icell]
        return self.iglob[o:o + self._ndofs_cell[icell]]


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
    """A scalar Lagrange FE space with a device-resident per-cell dof map."""

    def __init__(self, fs: FunctionSpace, mesh: Mesh, isContinuous: bool = True):
        self.fs = fs
        self.mesh = mesh
        self.continuous = bool(isContinuous)
        self.dhl = DofHandler(mesh, fs, 1, self.continuous)
        self._handle = None

    def get_ndofs(self) -> int:
        return self.dhl.ndofs_tot

    def get_dofs(self, icell: int) -> np.ndarray:
        return self.dhl.get_dof(icell)

    @property
    def handle(self):
        """bh_space on the mesh's backend, numbered from this space's DofHandler."""
        if self._handle is None:
            h = C.c_void_p()
            m = self.mesh.handle
            L.check(L.lib().bh_space_create_from_dofs(
                m, self.fs.degree, int(self.continuous), self.dhl.ndofs_tot,
                self.dhl.iglob.ctypes.data, self.dhl.offset.ctypes.data, C.byref(h)),
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
                 isContinuous: bool | None = None):
        if isinstance(kind, dict):  # TrialFESpace(fs, mesh, Dict(...))
            kind, dirichlet = "continuous", kind
        cont = (kind != "discontinuous") if isContinuous is None else bool(isContinuous)
        super().__init__(fs, mesh, cont)
        self.dirichlet = dict(dirichlet or {})


class TestFESpace:
    """TestFESpace(U) (fespace.jl:414): shares U's numbering."""

    __test__ = False  # not a pytest class

    def __init__(self, U: SingleFESpace):
        self.U = U

    def __getattr__(self, name):
        return getattr(self.U, name)
