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
from .forms import Placeholder, lower_bilinear, lower_dg_transport, lower_linear

_COEF = {"const": L.COEF_CONST, "cell": L.COEF_CELL, "qp": L.COEF_QP}
_FORM = {"stiffness": L.FORM_STIFFNESS, "mass": L.FORM_MASS, "source": L.FORM_SOURCE}


def _integrals(form_value):
    return form_value.terms()


def _single_measure(items):
    m = items[0].measure
    for it in items[1:]:
        if it.measure.degree != m.degree or it.measure.domain.mesh is not m.domain.mesh:
            raise NotImplementedError("all integrals of a form must share one Measure")
    return m


class BilinearPlan:
    """Pattern + terms of one bilinear form, reusable across re-assemblies
    (time stepping, Newton).  The pattern is built once; `assemble` writes the
    CSC values into HBM (device pointer) or into a host array."""

    def __init__(self, a, U, V, mode: str = "fp64"):
        U = getattr(U, "U", U)
        V = getattr(V, "U", V)
        if U is not V:
            raise NotImplementedError("assembly with distinct trial and test spaces")
        self.U, self.V = U, V
        self.mode = L.MODE[mode]
        items = _integrals(a(Placeholder("trial", U), Placeholder("test", V)))
        self.measure = _single_measure(items)
        if self.measure.domain.mesh is not V.mesh:
            raise ValueError("the measure's mesh differs from the space's mesh")
        self._keep = []
        terms = []
        for it in items:
            form, kind, data, scale = lower_bilinear(it.expr)
            t = L.BhTerm(_FORM[form], _COEF[kind], None, scale)
            if kind == "cell":
                arr = np.ascontiguousarray(data, dtype=np.float64)
                if arr.size != V.mesh.ncells:
                    raise ValueError("MeshCellData size differs from the number of cells")
                self._keep.append(arr)
                t.coef = arr.ctypes.data
            terms.append(t)
        self.terms = (L.BhTerm * len(terms))(*terms)
        self.nterms = len(terms)
        ctx = V.mesh.backend.ctx
        h = C.c_void_p()
        L.check(L.lib().bh_pattern_build(U.handle, V.handle, C.byref(h)), ctx, "bh_pattern_build")
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
    """assemble_bilinear(a, U, V; T=Float64) (assembler.jl:172-186)."""
    if T is not np.float64:
        raise NotImplementedError("only Float64 output is supported")
    plan = BilinearPlan(a, U, V, mode)
    return plan.to_csc(plan.assemble_host())


def _linear_terms(l, V):
    items = _integrals(l(Placeholder("test", V)))
    measure = _single_measure(items)
    mesh = V.mesh
    keep, terms = [], []
    xq = None
    for it in items:
        form, kind, data, scale = lower_linear(it.expr)
        t = L.BhTerm(_FORM[form], _COEF[kind], None, scale)
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
        terms.append(t)
    return measure, keep, (L.BhTerm * len(terms))(*terms), len(terms)


def assemble_linear_(b: np.ndarray, l, V, mode: str = "fp64") -> None:
    """assemble_linear!(b, l, V) (assembler.jl:465-473): adds l(v) into b."""
    V = getattr(V, "U", V)
    if b.dtype != np.float64 or not b.flags.c_contiguous or b.size != V.get_ndofs():
        raise ValueError("b must be a contiguous float64 vector of ndofs entries")
    measure, keep, terms, n = _linear_terms(l, V)
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
    V = getattr(V, "U", V)
    items = _integrals(l(Placeholder("test", V)))
    dg = lower_dg_transport(items)
    if dg is not None:
        c, u, degree = dg
        return _dg_plan(V, c, degree, mode).residual(u.values)
    b = np.zeros(V.get_ndofs(), dtype=np.float64)
    assemble_linear_(b, l, V, mode)
    return b


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
