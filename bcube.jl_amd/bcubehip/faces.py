"""Bilinear and linear forms with face integrals (interior and boundary faces)
on Quad4 meshes: the symmetric interior-penalty Poisson forms of
benchmark/bench_assemble.jl:33-96,

    a_Ω(u,v)  = ∫(∇(v)⋅∇(u))dΩ
    a_Γ(u,v)  = ∫(-jump(v,nΓ)⋅avg(∇(u)) - avg(∇(v))⋅jump(u,nΓ) + γ/h jump(v,nΓ)⋅jump(u,nΓ))dΓ
    a_Γb(u,v) = ∫(fa_Γb ∘ map(side⁻, (u, ∇(u), v, ∇(v), nΓb)))dΓb
    l_Γb(v)   = ∫(fl_Γb ∘ map(side⁻, (v, ∇(v), nΓb, g)))dΓb

assembled by the reference through assemble_bilinear / assemble_linear over
InteriorFaceDomain and BoundaryFaceDomain measures (assembler.jl:172-310,
448-520).  The cell terms run through the element kernels, the face terms
through one workgroup per face, and one row gather sums everything into a
single CSC.  Face integrands are lowered symbolically (forms.lower_face_*).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import scipy.sparse as sp

from . import _lib as L
from .forms import (BoundaryFaceDomain, CellDomain, InteriorFaceDomain, Placeholder, _split_sum,
                    lower_bilinear, lower_face_bilinear, lower_face_linear)


def has_faces(items) -> bool:
    return any(isinstance(it.measure.domain, (InteriorFaceDomain, BoundaryFaceDomain))
               for it in items)


def _split(items):
    vol = [it for it in items if isinstance(it.measure.domain, CellDomain)]
    inner = [it for it in items if isinstance(it.measure.domain, InteriorFaceDomain)]
    bnd = [it for it in items if isinstance(it.measure.domain, BoundaryFaceDomain)]
    return vol, inner, bnd


def _one_degree(items, what):
    degs = {it.measure.degree for it in items}
    if len(degs) > 1:
        raise NotImplementedError(f"all {what} integrals must use one quadrature degree")
    return degs.pop() if degs else None


def _check_mesh(items, mesh):
    for it in items:
        if it.measure.domain.mesh is not mesh:
            raise ValueError("a measure's mesh differs from the space's mesh")
        d = it.measure.domain
        if isinstance(d, BoundaryFaceDomain) and d.bc is not None:
            raise NotImplementedError(
                "face terms over a periodic BoundaryFaceDomain (only the Euler residual "
                "pairs periodic faces)")


class FaceBilinearPlan:
    """The CSC structure and face/cell terms of one bilinear form with face
    integrals, reusable across re-assemblies (bh_dgb_*)."""

    def __init__(self, a, U, V, face_degree: int | None = None):
        from .assembler import _bilinear_bh_terms
        U = getattr(U, "U", U)
        V = getattr(V, "U", V)
        if U is not V:
            raise NotImplementedError("assembly with distinct trial and test spaces")
        self.V = V
        items = a(Placeholder("trial", U), Placeholder("test", V)).terms()
        _check_mesh(items, V.mesh)
        vol, inner, bnd = _split(items)
        fdeg = _one_degree(inner + bnd, "face")
        vdeg = _one_degree(vol, "cell")
        self.face_degree = fdeg if fdeg is not None else (face_degree or 1)
        self.vol_degree = vdeg if vdeg is not None else 0
        lowered = [lower_bilinear(t) for it in vol for t in _split_sum(it.expr)]
        self._keep, vterms, nvol = _bilinear_bh_terms(lowered, V.mesh) if lowered else ([], None, 0)
        self.interior = self.boundary = None
        if inner:
            self.interior = np.ascontiguousarray(
                sum(lower_face_bilinear(it.expr, False) for it in inner))
        if bnd:
            self.boundary = np.ascontiguousarray(
                sum(lower_face_bilinear(it.expr, True) for it in bnd))
        self.ctx = V.mesh.backend.ctx
        h = C.c_void_p()
        ptr = lambda x: None if x is None else x.ctypes.data  # noqa: E731
        L.check(L.lib().bh_dgb_create(V.handle, self.face_degree, nvol, vterms, self.vol_degree,
                                      ptr(self.interior), ptr(self.boundary), C.byref(h)),
                self.ctx, "bh_dgb_create")
        self.handle = h
        n, nnz, ni, nb, nq = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int64(), C.c_int32()
        L.check(L.lib().bh_dgb_info(h, C.byref(n), C.byref(nnz), C.byref(ni), C.byref(nb),
                                    C.byref(nq)), self.ctx, "bh_dgb_info")
        self.n, self.nnz = n.value, nnz.value
        self.ninterior, self.nboundary, self.nq = ni.value, nb.value, nq.value
        self._colptr = None

    def csc_structure(self):
        if self._colptr is None:
            colptr = np.zeros(self.n + 1, dtype=np.int64)
            rowval = np.zeros(self.nnz, dtype=np.int64)
            L.check(L.lib().bh_dgb_export_csc(self.handle, colptr.ctypes.data, rowval.ctypes.data,
                                              0), self.ctx, "bh_dgb_export_csc")
            self._colptr, self._rowval = colptr, rowval
        return self._colptr, self._rowval

    def assemble_host(self, out: np.ndarray | None = None) -> np.ndarray:
        if out is None:
            out = np.empty(self.nnz, dtype=np.float64)
        if out.dtype != np.float64 or not out.flags.c_contiguous or out.size != self.nnz:
            raise ValueError("out must be a contiguous float64 vector of nnz entries")
        L.check(L.lib().bh_dgb_assemble(self.handle, out.ctypes.data, L.HOST), self.ctx,
                "bh_dgb_assemble")
        return out

    def assemble_device(self, nzval_ptr: int):
        """Enqueue assembly into device memory (nnz FP64 values)."""
        L.check(L.lib().bh_dgb_assemble(self.handle, nzval_ptr, L.DEVICE), self.ctx,
                "bh_dgb_assemble")

    def to_csc(self, nzval: np.ndarray) -> sp.csc_matrix:
        colptr, rowval = self.csc_structure()
        return sp.csc_matrix((nzval, rowval, colptr), shape=(self.n, self.n))

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            L.lib().bh_dgb_destroy(h)
            self.handle = None


class BoundaryFaces:
    """The boundary faces of a space's mesh with one quadrature rule: their
    quadrature points and the boundary linear forms (bh_dgb_boundary_*)."""

    def __init__(self, V, degree: int):
        self.V = V
        self.ctx = V.mesh.backend.ctx
        h = C.c_void_p()
        L.check(L.lib().bh_dgb_create(V.handle, degree, 0, None, 0, None, None, C.byref(h)),
                self.ctx, "bh_dgb_create")
        self.handle = h
        n, nnz, ni, nb, nq = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int64(), C.c_int32()
        L.check(L.lib().bh_dgb_info(h, C.byref(n), C.byref(nnz), C.byref(ni), C.byref(nb),
                                    C.byref(nq)), self.ctx, "bh_dgb_info")
        self.nboundary, self.nq = nb.value, nq.value
        self._xq = None

    def points(self) -> np.ndarray:
        """Physical quadrature points of the boundary faces, (nboundary, nq, 2)."""
        if self._xq is None:
            xq = np.zeros(self.nboundary * self.nq * 2, dtype=np.float64)
            L.check(L.lib().bh_dgb_boundary_points(self.handle, xq.ctypes.data, L.HOST),
                    self.ctx, "bh_dgb_boundary_points")
            self._xq = xq.reshape(self.nboundary, self.nq, 2)
        return self._xq

    def add_linear(self, b: np.ndarray, c0: float, c1: float, g=None) -> None:
        """b += Σ_faces ∫ (c0 v + c1 n⋅∇v) g dΓ, g a PhysicalFunction or None (g = 1)."""
        if b.dtype != np.float64 or not b.flags.c_contiguous or b.size != self.V.get_ndofs():
            raise ValueError("b must be a contiguous float64 vector of ndofs entries")
        if g is None:
            vals = np.ones(self.nboundary * self.nq, dtype=np.float64)
        else:
            vals = np.ascontiguousarray(
                np.asarray(g.f(self.points().reshape(-1, 2)), dtype=np.float64).reshape(-1))
            if vals.size != self.nboundary * self.nq:
                raise ValueError("the PhysicalFunction returned the wrong number of values")
        L.check(L.lib().bh_dgb_boundary_linear(self.handle, c0, c1, vals.ctypes.data,
                                               b.ctypes.data, L.HOST), self.ctx,
                "bh_dgb_boundary_linear")

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            L.lib().bh_dgb_destroy(h)
            self.handle = None


def boundary_face_set(V, degree: int) -> BoundaryFaces:
    """The space's BoundaryFaces for one quadrature degree, built once."""
    cache = V.__dict__.setdefault("_boundary_faces", {})
    if degree not in cache:
        cache[degree] = BoundaryFaces(V, degree)
    return cache[degree]


def assemble_linear_faces_(b: np.ndarray, items, V) -> list:
    """Adds the BoundaryFaceDomain integrals of a linear form into b and
    returns the other integrals (InteriorFaceDomain linear terms raise)."""
    _check_mesh(items, V.mesh)
    vol, inner, bnd = _split(items)
    if inner:
        raise NotImplementedError("linear forms over interior faces")
    for it in bnd:
        faces = boundary_face_set(V, it.measure.degree)
        for c0, c1, g in lower_face_linear(it.expr):
            faces.add_linear(b, c0, c1, g)
    return vol
