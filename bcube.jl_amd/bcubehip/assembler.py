"""assemble_bilinear / assemble_linear on the HIP backend.

Mirrors src/assembler/assembler.jl: assemble_bilinear (:172-186) returns the
SparseMatrixCSC{Float64,Int64} analogue (scipy.sparse.csc_matrix with int64
indices), assemble_linear (:448-458) a float64 vector, assemble_linear_ is the
in-place assemble_linear! (:465-473).  Every call runs on the GPU through
libbcubehip; unsupported integrands raise.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import scipy.sparse as sp

from . import _lib as L
from .faces import FaceBilinearPlan, assemble_linear_faces_, has_faces
from .forms import (CellDomain, Placeholder, _split_sum, lower_bilinear, lower_dg_transport, lower_linear,
                    lower_bilinear_terms, lower_linear_terms,
                    lower_multi_bilinear, lower_multi_linear)

_COEF = {"const": L.COEF_CONST, "cell": L.COEF_CELL, "qp": L.COEF_QP, "fe": L.COEF_QP}
_FORM = {"stiffness": L.FORM_STIFFNESS, "mass": L.FORM_MASS, "source": L.FORM_SOURCE}


def _integrals(form_value):
    return form_value.terms()


def _cells_only(items, what):
    """Vector and multi-field assembly integrates over cells only."""
    for it in items:
        if not isinstance(it.measure.domain, CellDomain):
            raise NotImplementedError(
                f"{what} over a vector space or MultiFESpace: only CellDomain measures")


def _single_measure(items):
    m = items[0].measure
    for it in items[1:]:
        if it.measure.degree != m.degree or it.measure.domain.mesh is not m.domain.mesh:
            raise NotImplementedError("all integrals of a form must share one Measure")
    return m


def _bilinear_bh_terms(lowered, mesh):
    keep, terms = [], []
    arrays = {}  # one array per coefficient object: the axis terms of (c⋅∇u) v share it
    for form, kind, data, scale in lowered:
        if isinstance(form, tuple):  # ("adv", a): ∫ η ∂u/∂x_a v
            if form[1] >= mesh.dim:
                raise NotImplementedError("the vector in (c⋅∇u) v has more entries than the "
                                          "mesh has dimensions")
            code = L.FORM_ADVECTION[form[1]]
        else:
            code = _FORM[form]
        t = L.BhTerm(code, _COEF[kind], None, scale)
        if kind == "cell":
            arr = arrays.get(id(data))
            if arr is None:
                arr = arrays[id(data)] = np.ascontiguousarray(data, dtype=np.float64)
            if arr.size != mesh.ncells:
                raise ValueError("MeshCellData size differs from the number of cells")
            keep.append(arr)
            t.coef = arr.ctypes.data
        terms.append(t)
    return keep, (L.BhTerm * len(terms))(*terms), len(terms)


def _is_multi(U) -> bool:
    from .fespace import MultiFESpace
    return isinstance(U, MultiFESpace) or getattr(getattr(U, "U", U), "size", 1) > 1


class Fields:
    """The fields of a vector space or MultiFESpace: every component of every
    space is a copy of one scalar space's dofs, numbered into the global
    system by maps[f] (0-based) -- the vector numbering of dofhandler.jl:37-156
    composed with get_mapping (fespace.jl:610-647)."""

    def __init__(self, U):
        from .fespace import MultiFESpace, _single
        if isinstance(U, MultiFESpace):
            self.spaces = [_single(s) for s in U.spaces]
            mappings = [U.get_mapping(i) - 1 for i in range(len(self.spaces))]
            self.ndofs = U.get_ndofs()
        else:
            self.spaces = [_single(U)]
            mappings = [np.arange(self.spaces[0].get_ndofs(), dtype=np.int64)]
            self.ndofs = self.spaces[0].get_ndofs()
        s0 = self.spaces[0]
        for sp in self.spaces[1:]:
            if (sp.mesh is not s0.mesh or sp.fs.degree != s0.fs.degree
                    or sp.continuous != s0.continuous):
                raise NotImplementedError(
                    "fields of different meshes, orders or continuity (rectangular blocks)")
        self.scalar = s0
        self.sizes = [sp.size for sp in self.spaces]
        self.first = np.concatenate([[0], np.cumsum(self.sizes)])[:-1].tolist()
        maps = []
        for sp, mp in zip(self.spaces, mappings):
            cm = sp.component_map
            for k in range(sp.size):
                maps.append(mp[cm[k]])
        self.maps = np.ascontiguousarray(np.stack(maps), dtype=np.int64)

    def field(self, i: int, k: int) -> int:
        return self.first[i] + k

    def placeholders(self, U, role):
        from .fespace import MultiFESpace
        if isinstance(U, MultiFESpace):
            return tuple(Placeholder(role, s, field=i) for i, s in enumerate(U.spaces))
        return Placeholder(role, U)


class MultiBilinearPlan:
    """assemble_bilinear over a vector space or MultiFESpace (assembler.jl:
    327-420): one scalar block per distinct coupled form, assembled by the
    element kernels into library-owned device buffers over the scalar pattern,
    then spread into the global CSC by bh_blocks_assemble."""

    def __init__(self, a, U, V, mode: str = "fp64"):
        self.mode = L.MODE[mode]
        fu, fv = Fields(U), Fields(V)
        if fu.sizes != fv.sizes or not np.array_equal(fu.maps, fv.maps):
            raise NotImplementedError("trial and test fields must be numbered alike")
        self.fields = fu
        sizes = fu.sizes
        items = _integrals(a(fu.placeholders(U, "trial"), fv.placeholders(V, "test")))
        _cells_only(items, "bilinear forms")
        self.measure = _single_measure(items)
        mesh = fu.scalar.mesh
        if self.measure.domain.mesh is not mesh:
            raise ValueError("the measure's mesh differs from the space's mesh")
        by_pair = lower_multi_bilinear(items, sizes, mesh.dim)
        self._keep, self.blocks = [], []
        pairs, blk = [], []
        seen = {}  # identical scalar forms share one block
        for (i, j), comp in by_pair.items():
            for ki in range(sizes[i]):
                for kj in range(sizes[j]):
                    lowered = comp.get((ki, kj))
                    b = -1  # a component pair the form does not couple: explicit zeros
                    if lowered is not None:
                        key = tuple((f, k, id(d) if d is not None else None, sc)
                                    for f, k, d, sc in lowered)
                        if key not in seen:
                            keep, terms, n = _bilinear_bh_terms(lowered, mesh)
                            self._keep.append((keep, lowered))
                            seen[key] = len(self.blocks)
                            self.blocks.append((terms, n))
                        b = seen[key]
                    pairs += [fu.field(i, ki), fu.field(j, kj)]
                    blk.append(b)
        if len(self.blocks) > 32:
            raise NotImplementedError("more than 32 coupled forms")
        self.block_of_pair = np.ascontiguousarray(blk, dtype=np.int8)
        self.pattern = pattern_of(fu.scalar)
        self.ctx = mesh.backend.ctx
        h = C.c_void_p()
        pr = np.ascontiguousarray(pairs, dtype=np.int32)
        L.check(L.lib().bh_blocks_create(self.pattern.handle, len(fu.maps), fu.maps.ctypes.data,
                                         fu.ndofs, len(blk), pr.ctypes.data, C.byref(h)),
                self.ctx, "bh_blocks_create")
        self.handle = h
        n, nnz = C.c_int64(), C.c_int64()
        L.check(L.lib().bh_blocks_info(h, C.byref(n), C.byref(nnz)), self.ctx, "bh_blocks_info")
        self.nrows = self.ncols = n.value
        self.nnz = nnz.value
        self._colptr = None

    def csc_structure(self):
        if self._colptr is None:
            colptr = np.zeros(self.ncols + 1, dtype=np.int64)
            rowval = np.zeros(self.nnz, dtype=np.int64)
            L.check(L.lib().bh_blocks_export_csc(self.handle, colptr.ctypes.data,
                                                 rowval.ctypes.data, 0), self.ctx,
                    "bh_blocks_export_csc")
            self._colptr, self._rowval = colptr, rowval
        return self._colptr, self._rowval

    def _blocks(self, mode, coef_ptrs=None):
        # each scalar block into the library's buffer k; per-cell coefficient
        # arrays are the plan's host arrays (staged by the library) unless
        # coef_ptrs[k] gives device pointers for block k, in term order
        md = self.mode if mode is None else L.MODE[mode]
        for k, (terms, n) in enumerate(self.blocks):
            where = L.HOST
            if coef_ptrs is not None and coef_ptrs[k] is not None:
                terms = (L.BhTerm * n)(*[L.BhTerm(t.form, t.coef_kind, p, t.scale)
                                         for t, p in zip(terms, coef_ptrs[k])])
                where = L.DEVICE
            L.check(L.lib().bh_blocks_assemble_block(self.handle, k, n, terms,
                                                     self.measure.degree, md, where),
                    self.ctx, "bh_blocks_assemble_block")

    def assemble_host(self, out: np.ndarray | None = None, mode: str | None = None) -> np.ndarray:
        if out is None:
            out = np.empty(self.nnz, dtype=np.float64)
        self._blocks(mode)
        L.check(L.lib().bh_blocks_assemble(self.handle, len(self.blocks), None,
                                           self.block_of_pair.ctypes.data, out.ctypes.data,
                                           L.HOST), self.ctx, "bh_blocks_assemble")
        return out

    def assemble_device(self, nzval_ptr: int, mode: str | None = None, coef_ptrs=None):
        """Enqueue the assembly into device memory; coef_ptrs: per block (in
        the order of `blocks`), None or the device pointers of its terms'
        coefficient arrays (None for a constant term)."""
        self._blocks(mode, coef_ptrs)
        L.check(L.lib().bh_blocks_assemble(self.handle, len(self.blocks), None,
                                           self.block_of_pair.ctypes.data, nzval_ptr, L.DEVICE),
                self.ctx, "bh_blocks_assemble")

    def to_csc(self, nzval: np.ndarray) -> sp.csc_matrix:
        colptr, rowval = self.csc_structure()
        A = sp.csc_matrix((nzval, rowval, colptr), shape=(self.nrows, self.ncols))
        A.has_sorted_indices = True
        return A

    def __del__(self):
        try:
            L.lib().bh_blocks_destroy(self.handle)
        except Exception:
            pass


def _fields_blocks(V):
    """Fields of V and a pair-free bh_blocks over its scalar pattern (for
    linear forms), built once per space object."""
    cache = V.__dict__.setdefault("_fields_blocks", {})
    if "b" not in cache:
        f = Fields(V)
        pat = pattern_of(f.scalar)
        h = C.c_void_p()
        ctx = f.scalar.mesh.backend.ctx
        L.check(L.lib().bh_blocks_create(pat.handle, len(f.maps), f.maps.ctypes.data, f.ndofs, 0,
                                         None, C.byref(h)), ctx, "bh_blocks_create")
        cache.update(b=h, f=f, pat=pat)
    return cache["f"], cache["b"]


def _assemble_linear_multi(b: np.ndarray, l, V, mode: str) -> None:
    f, h = _fields_blocks(V)
    items = _integrals(l(f.placeholders(V, "test")))
    _cells_only(items, "linear forms")
    measure = _single_measure(items)
    mesh = f.scalar.mesh
    ctx = mesh.backend.ctx
    tmp = np.empty(f.scalar.get_scalar_ndofs(), dtype=np.float64)
    for (i, k), lowered in lower_multi_linear(items, f.sizes).items():
        keep, terms, n = _linear_bh_terms(lowered, measure, mesh)
        tmp[:] = 0.0
        L.check(L.lib().bh_assemble_linear(f.scalar.handle, n, terms, measure.degree,
                                           L.MODE[mode], tmp.ctypes.data, L.HOST), ctx,
                "bh_assemble_linear")
        L.check(L.lib().bh_blocks_add_field(h, f.field(i, k), tmp.ctypes.data, b.ctypes.data,
                                            L.HOST), ctx, "bh_blocks_add_field")


class BilinearPlan:
    """Pattern + terms of one bilinear form, reusable across re-assemblies
    (time stepping, Newton).  The pattern is built once; `assemble` writes the
    CSC values into HBM (device pointer) or into a host array."""

    def __init__(self, a, U, V, mode: str = "fp64", build: str = "host"):
        """build="device" builds the pattern by kernels (bh_pattern_build_device,
        the same structure), for meshes that change."""
        U = getattr(U, "U", U)
        V = getattr(V, "U", V)
        if U is not V:
            # distinct trial and test spaces: one mesh, scalar, FP64
            # (an ndofs(V) x ndofs(U) matrix)
            if U.mesh is not V.mesh:
                raise NotImplementedError("trial and test spaces on different meshes")
            if getattr(U, "size", 1) != 1 or getattr(V, "size", 1) != 1:
                raise NotImplementedError("distinct trial and test spaces: scalar spaces only")
            if build != "host":
                raise NotImplementedError("distinct trial and test spaces: host pattern build")
        self.U, self.V = U, V
        self.mode = L.MODE[mode]
        items = _integrals(a(Placeholder("trial", U), Placeholder("test", V)))
        self.measure = _single_measure(items)
        if self.measure.domain.mesh is not V.mesh:
            raise ValueError("the measure's mesh differs from the space's mesh")
        lowered = [x for it in items for t in _split_sum(it.expr)
                   for x in lower_bilinear_terms(t, V.mesh.dim)]
        self._keep, self.terms, self.nterms = _bilinear_bh_terms(lowered, V.mesh)
        ctx = V.mesh.backend.ctx
        h = C.c_void_p()
        fn = {"host": "bh_pattern_build", "device": "bh_pattern_build_device"}[build]
        L.check(getattr(L.lib(), fn)(U.handle, V.handle, C.byref(h)), ctx, fn)
        self.handle = h
        nr, nc, nnz, ncol = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int32()
        L.check(L.lib().bh_pattern_info(h, C.byref(nr), C.byref(nc), C.byref(nnz), C.byref(ncol)),
                ctx, "bh_pattern_info")
        self.nrows, self.ncols, self.nnz, self.max_row = nr.value, nc.value, nnz.value, ncol.value
        self.ctx = ctx
        self._colptr = None

    def csc_structure(self):
        if self._colptr is None:
            colptr = np.zeros(self.ncols + 1, dtype=np.int64)
            rowval = np.zeros(self.nnz, dtype=np.int64)
            L.check(L.lib().bh_pattern_export_csc(self.handle, colptr.ctypes.data,
                                                  rowval.ctypes.data, 0), self.ctx,
                    "bh_pattern_export_csc")
            self._colptr, self._rowval = colptr, rowval
        return self._colptr, self._rowval

    def assemble_host(self, out: np.ndarray | None = None, mode: str | None = None) -> np.ndarray:
        if out is None:
            out = np.empty(self.nnz, dtype=np.float64)
        md = self.mode if mode is None else L.MODE[mode]
        L.check(L.lib().bh_assemble_bilinear(self.handle, self.nterms, self.terms,
                                             self.measure.degree, md, out.ctypes.data,
                                             L.HOST), self.ctx, "bh_assemble_bilinear")
        return out

    def assemble_device(self, nzval_ptr: int | None = None, coef_ptrs=None,
                        mode: str | None = None):
        """Enqueue assembly into device memory (nzval_ptr, or the pattern's own
        buffer if None).  coef_ptrs: optional device pointers replacing the
        per-cell coefficient arrays, in term order (None keeps a term as is).
        mode overrides the plan's arithmetic for this call."""
        terms = self.terms
        if coef_ptrs is not None:
            terms = (L.BhTerm * self.nterms)(*[L.BhTerm(t.form, t.coef_kind,
                                                        p if p is not None else None, t.scale)
                                               for t, p in zip(self.terms, coef_ptrs)])
        md = self.mode if mode is None else L.MODE[mode]
        L.check(L.lib().bh_assemble_bilinear(self.handle, self.nterms, terms, self.measure.degree,
                                             md, nzval_ptr, L.DEVICE), self.ctx,
                "bh_assemble_bilinear")

    def set_offset_templates(self, enable: bool):
        """Row gather offsets from the pattern's templates (default, where it has
        them) or per entry (bh_pattern_set_offset_templates); same values."""
        L.check(L.lib().bh_pattern_set_offset_templates(self.handle, int(enable)), self.ctx,
                "bh_pattern_set_offset_templates")

    def offset_templates(self) -> int:
        """Number of offset templates the row gather uses (0: per-entry offsets)."""
        n = L.lib().bh_pattern_offset_templates(self.handle)
        if n < 0:
            L.check(n, self.ctx, "bh_pattern_offset_templates")
        return n

    def tuning(self, mode: str | None = None):
        """Launch shapes the autotuner chose for this plan and mode, or None
        before tuning: element-kernel waves per block target, gather lanes per
        row and rows per lane group (0 = default), and their times in ms."""
        md = self.mode if mode is None else L.MODE[mode]
        shape = np.zeros(3, dtype=np.int32)
        ms = np.zeros(2, dtype=np.float32)
        rc = L.lib().bh_pattern_tuning(self.handle, md, shape.ctypes.data, ms.ctypes.data)
        if rc < 0:
            L.check(rc, self.ctx, "bh_pattern_tuning")
        if rc == 0:
            return None
        return {"elem_waves": int(shape[0]), "gather_lanes": int(shape[1]),
                "gather_rows": int(shape[2]), "elem_ms": float(ms[0]), "gather_ms": float(ms[1])}

    def set_tuning(self, elem_waves: int = 0, gather_lanes: int = 0, gather_rows: int = 0,
                   mode: str | None = None):
        """Fix the launch shapes for a mode (0 = default); raises BcubeHipError
        for a gather shape that is not a candidate of this pattern."""
        md = self.mode if mode is None else L.MODE[mode]
        shape = np.array([elem_waves, gather_lanes, gather_rows], dtype=np.int32)
        L.check(L.lib().bh_pattern_set_tuning(self.handle, md, shape.ctypes.data), self.ctx,
                "bh_pattern_set_tuning")

    def values_device_ptr(self) -> int:
        p = C.c_void_p()
        L.check(L.lib().bh_pattern_values_device(self.handle, C.byref(p)), self.ctx,
                "bh_pattern_values_device")
        return p.value

    def onepass(self):
        """The one-pass structured Q2 kernel's analysis of this pattern
        (bh_pattern_onepass): {"ok", "nx", "ny", "nz", "templates",
        "layers_per_group", "why"}."""
        info = np.zeros(6, dtype=np.int32)
        L.check(L.lib().bh_pattern_onepass(self.handle, info.ctypes.data), self.ctx,
                "bh_pattern_onepass")
        why = "" if info[0] else L.lib().bh_last_error(self.ctx).decode()
        return {"ok": bool(info[0]), "nx": int(info[1]), "ny": int(info[2]), "nz": int(info[3]),
                "templates": int(info[4]), "layers_per_group": int(info[5]), "why": why}

    def to_csc(self, nzval: np.ndarray) -> sp.csc_matrix:
        colptr, rowval = self.csc_structure()
        A = sp.csc_matrix((nzval, rowval, colptr), shape=(self.nrows, self.ncols))
        A.has_sorted_indices = True
        return A

    def __del__(self):
        try:
            L.lib().bh_pattern_destroy(self.handle)
        except Exception:
            pass


class Pattern:
    """The CSR pattern of (U, U) alone (bh_pattern_build), for edits of
    assembled matrices such as Dirichlet conditions."""

    def __init__(self, U):
        self.U = getattr(U, "U", U)
        self.ctx = self.U.mesh.backend.ctx
        h = C.c_void_p()
        L.check(L.lib().bh_pattern_build(self.U.handle, self.U.handle, C.byref(h)), self.ctx,
                "bh_pattern_build")
        self.handle = h
        nr, nc, nnz, mr = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int32()
        L.check(L.lib().bh_pattern_info(h, C.byref(nr), C.byref(nc), C.byref(nnz), C.byref(mr)),
                self.ctx, "bh_pattern_info")
        self.nnz = nnz.value

    def __del__(self):
        try:
            L.lib().bh_pattern_destroy(self.handle)
        except Exception:
            pass


def pattern_of(U) -> Pattern:
    """The space's (U, U) pattern, built once."""
    U = getattr(U, "U", U)
    p = U.__dict__.get("_pattern")
    if p is None:
        p = U.__dict__["_pattern"] = Pattern(U)
    return p


def assemble_bilinear(a, U, V, T=np.float64, mode: str = "fp64") -> sp.csc_matrix:
    """assemble_bilinear(a, U, V; T=Float64) (assembler.jl:172-186); over vector
    spaces and MultiFESpaces (assembler.jl:327-420) a, U and V are those of the
    reference, a receiving tuples for MultiFESpaces."""
    if T is not np.float64:
        raise NotImplementedError("only Float64 output is supported")
    if not (_is_multi(U) or _is_multi(V)):
        Us, Vs = getattr(U, "U", U), getattr(V, "U", V)
        if has_faces(_integrals(a(Placeholder("trial", Us), Placeholder("test", Vs)))):
            if mode != "fp64":
                raise NotImplementedError("face terms are assembled in FP64")
            plan = FaceBilinearPlan(a, Us, Vs)
            return plan.to_csc(plan.assemble_host())
    plan = (MultiBilinearPlan if _is_multi(U) or _is_multi(V) else BilinearPlan)(a, U, V, mode)
    return plan.to_csc(plan.assemble_host())


def _linear_terms(l, V, items=None):
    if items is None:
        items = _integrals(l(Placeholder("test", V)))
    measure = _single_measure(items)
    return (measure,) + _linear_bh_terms([t for it in items for t in lower_linear_terms(it.expr)],
                                         measure,
                                         V.mesh)


def _linear_bh_terms(lowered, measure, mesh):
    keep, terms = [], []
    xq = None
    for form, kind, data, scale in lowered:
        if isinstance(form, tuple):  # ("grad", a): ∫ f ∂v/∂x_a
            if form[1] >= mesh.dim:
                raise NotImplementedError("the vector in (f c)⋅∇v has more entries than the mesh "
                                          "has dimensions")
            code = L.FORM_GRAD_SOURCE[form[1]]
        else:
            code = _FORM[form]
        t = L.BhTerm(code, _COEF[kind], None, scale)
        if kind == "cell":
            arr = np.ascontiguousarray(data, dtype=np.float64)
            keep.append(arr)
            t.coef = arr.ctypes.data
        elif kind == "qp":
            if xq is None:
                xq = quadrature_points(mesh, measure.degree)
            vals = np.ascontiguousarray(np.asarray(data(xq.reshape(-1, mesh.dim)),
                                                   dtype=np.float64).reshape(-1))
            keep.append(vals)
            t.coef = vals.ctypes.data
        elif kind == "fe":
            if data.space.mesh is not mesh:
                raise NotImplementedError("an FE function of another mesh")
            vals = fe_values_at_points(data, measure.degree).reshape(-1)
            keep.append(vals)
            t.coef = vals.ctypes.data
        terms.append(t)
    return keep, (L.BhTerm * len(terms))(*terms), len(terms)


def assemble_linear_(b: np.ndarray, l, V, mode: str = "fp64") -> None:
    """assemble_linear!(b, l, V) (assembler.jl:465-473): adds l(v) into b."""
    if not _is_multi(V):
        V = getattr(V, "U", V)
    if b.dtype != np.float64 or not b.flags.c_contiguous or b.size != V.get_ndofs():
        raise ValueError("b must be a contiguous float64 vector of ndofs entries")
    if _is_multi(V):
        from .euler import euler_plan, lower_euler_dg
        from .fespace import MultiFESpace
        if isinstance(V, MultiFESpace):
            f, _ = _fields_blocks(V)
            eu = lower_euler_dg(_integrals(l(f.placeholders(V, "test"))), V)
            if eu is not None:
                # the COVO Euler DG residual (benchmark/covo.jl): one kernel
                gamma, u, degree, periodic, terms, fcoef = eu
                if mode != "fp64":
                    raise NotImplementedError("the Euler residual is assembled in FP64")
                euler_plan(V, gamma, degree, periodic).residual(u.values, b, terms, fcoef)
                return
        _assemble_linear_multi(b, l, V, mode)
        return
    items = _integrals(l(Placeholder("test", V)))
    if has_faces(items):
        # boundary-face integrals by the face kernels, the cell integrals below
        items = assemble_linear_faces_(b, items, V)
        if not items:
            return
    measure, keep, terms, n = _linear_terms(l, V, items)
    ctx = V.mesh.backend.ctx
    L.check(L.lib().bh_assemble_linear(V.handle, n, terms, measure.degree, L.MODE[mode],
                                       b.ctypes.data, L.HOST), ctx, "bh_assemble_linear")


def _dg_plan(V, c, degree, mode):
    from .dg import DGTransport
    cache = V.__dict__.setdefault("_dg_plans", {})
    key = (tuple(float(x) for x in c), int(degree), mode)
    if key not in cache:
        cache[key] = DGTransport(V, c, degree, mode)
    return cache[key]


def assemble_linear(l, V, T=np.float64, mode: str = "fp64") -> np.ndarray:
    """assemble_linear(l, V; T=Float64) (assembler.jl:448-458).

    The upwind DG transport form (see bcubehip.dg) goes to the DG kernels; its
    FEFunction may carry several fields, (nfields, ndofs), and the result has
    the same shape (FP32 in mode "fp32")."""
    if _is_multi(V):
        b = np.zeros(V.get_ndofs(), dtype=np.float64)
        assemble_linear_(b, l, V, mode)
        return b
    V = getattr(V, "U", V)
    items = _integrals(l(Placeholder("test", V)))
    dg = lower_dg_transport(items)
    if dg is not None:
        c, u, degree = dg
        return _dg_plan(V, c, degree, mode).residual(u.values)
    b = np.zeros(V.get_ndofs(), dtype=np.float64)
    assemble_linear_(b, l, V, mode)
    return b


def fe_values_at_points(u, degree: int) -> np.ndarray:
    """Values of the FEFunction u at the quadrature points of Measure(Ω, degree),
    (ncells, nq), computed on the GPU (bh_fe_eval_qp)."""
    U = u.space
    mesh = U.mesh
    ctx = mesh.backend.ctx
    nq = C.c_int32()
    L.check(L.lib().bh_quadrature_points(mesh.handle, degree, None, L.HOST, C.byref(nq)), ctx,
            "bh_quadrature_points")
    vals = np.ascontiguousarray(u.values, dtype=np.float64)
    if vals.shape != (U.get_ndofs(),):
        raise ValueError("FE function values do not match its space")
    out = np.zeros(mesh.ncells * nq.value, dtype=np.float64)
    L.check(L.lib().bh_fe_eval_qp(U.handle, degree, vals.ctypes.data, out.ctypes.data, L.HOST),
            ctx, "bh_fe_eval_qp")
    return out.reshape(mesh.ncells, nq.value)


def quadrature_points(mesh, degree: int) -> np.ndarray:
    """Physical quadrature points, (ncells, nq, dim), computed on the GPU."""
    nq = C.c_int32()
    ctx = mesh.backend.ctx
    L.check(L.lib().bh_quadrature_points(mesh.handle, degree, None, L.HOST, C.byref(nq)), ctx,
            "bh_quadrature_points")
    out = np.zeros(mesh.ncells * nq.value * mesh.dim, dtype=np.float64)
    got = C.c_int32()
    L.check(L.lib().bh_quadrature_points(mesh.handle, degree, out.ctypes.data, L.HOST,
                                         C.byref(got)), ctx, "bh_quadrature_points")
    return out.reshape(mesh.ncells, got.value, mesh.dim)
