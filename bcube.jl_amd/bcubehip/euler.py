"""The Euler DG residual of the reference's COVO benchmark (benchmark/covo.jl,
driver_bench_covo.jl) on the HIP backend.

    u  = FEFunction(U), U = MultiFESpace(U_sca, U_vec, U_sca, U_sca)  (DG order 1
         on Quad4, Tri3 or mixed cells)
    l(v) = ∫(flux_Ω(u, v))dΩ - ∫(flux_Γ(u, v, nΓ))dΓ
           - ∫(flux_Γ(u, v, nx))dΓ_perio_x - ∫(flux_Γ(u, v, ny))dΓ_perio_y
    flux_Ω(u, v)    = compose(EulerFlux(γ), u, tuple(map(grad, v)))
    flux_Γ(u, v, n) = compose(RoeFlux(γ), side_n(u), side_p(u), jump(v), side_n(n))

assemble_linear / assemble_linear_ recognise this form (any subset of its
integrals) and run it as one kernel (bh_euler_*); EulerDG also gives the
explicit step u + Δt M⁻¹ b(u) of covo.jl:14-33.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .forms import (BoundaryFaceDomain, CellDomain, Compose, Const, EulerFlux, FaceNormal,
                    FEFunction, Grad, InteriorFaceDomain, Jump, Mul, Placeholder, RoeFlux, Side,
                    _factors)

_SQ_F2N = [(0, 1), (1, 2), (2, 3), (3, 0)]  # faces2nodes of the square (shape.jl:254)


def periodic_face_pairs(mesh, bc) -> np.ndarray:
    """_compute_periodicity (domain.jl:269-397) and the periodic FaceInfo
    (domain.jl:845-890) on a Quad4/Tri3 mesh: for each slave face (face
    order) the master face whose centre maps onto it; side⁻ is the slave cell,
    side⁺ the master cell.  Rows [cell⁻, side⁻, cell⁺, side⁺, flip], flip = 1
    when the master side's first node maps onto the slave side's second node."""
    from .dirichlet import _SIDES, boundary_faces
    if np.any((mesh.celltypes != L.CELLTYPE["Quad4"]) & (mesh.celltypes != L.CELLTYPE["Tri3"])):
        raise NotImplementedError("periodic faces on Quad4 and Tri3 meshes")
    tags = boundary_faces(mesh)
    for lab in bc.labels_master + bc.labels_slave:
        if lab not in tags:
            raise ValueError(f"unknown boundary tag {lab!r}")
    X = mesh.nodes()
    master = [f for lab in bc.labels_master for f in tags[lab]]
    slave = [f for lab in bc.labels_slave for f in tags[lab]]

    def cnodes(c):
        return mesh.c2n[mesh.c2n_ptr[c]:mesh.c2n_ptr[c + 1]]

    def fnodes(c, s):
        nodes = cnodes(c)
        return [int(nodes[k]) for k in _SIDES[int(mesh.celltypes[c])][s]]

    mc = np.array([X[fnodes(c, s)].mean(axis=0) for c, s in master]).reshape(-1, 2)
    mimg = np.array([bc.transformation(x) for x in mc]).reshape(-1, 2)
    out = []
    for ci, si in slave:
        ni = fnodes(ci, si)
        Mi = X[ni].mean(axis=0)
        dx = np.linalg.norm(X[cnodes(ci)].mean(axis=0) - Mi) * np.sqrt(mesh.dim)
        d = np.linalg.norm(mimg - Mi, axis=1) if len(master) else np.array([])
        hit = np.flatnonzero(d <= bc.tol * dx)
        if hit.size == 0:
            raise ValueError(f"no periodic counterpart for the face of cell {ci}, side {si}")
        cj, sj = master[int(hit[0])]
        nj = fnodes(cj, sj)
        first = bc.transformation(X[nj[0]])
        flip = 0 if np.linalg.norm(first - X[ni[0]]) <= bc.tol * dx else 1
        out.append([ci, si, cj, sj, flip])
    return np.ascontiguousarray(np.array(out, dtype=np.int32).reshape(-1, 5))


class EulerDG:
    """The COVO residual on one MultiFESpace: side table, component maps and
    rule built once (bh_euler_create); residual(u) adds l(v) into b, step(u,
    dt) returns u + dt M⁻¹ b(u)."""

    def __init__(self, V, gamma: float, degree: int, periodic_domains=()):
        from .assembler import Fields
        f = Fields(V)
        if f.sizes != [1, 2, 1, 1]:
            raise NotImplementedError("the Euler residual needs MultiFESpace(ρ, ρu (size 2), ρE, ϕ)")
        sc = f.scalar
        if sc.continuous or sc.fs.degree != 1:
            raise NotImplementedError("the Euler residual needs discontinuous order-1 spaces")
        self.fields = f
        self.ndofs = f.ndofs
        pairs = [d.periodic_faces() for d in periodic_domains]
        per = (np.ascontiguousarray(np.concatenate(pairs), dtype=np.int32) if pairs
               else np.zeros((0, 5), dtype=np.int32))
        self._per = per
        self.ctx = sc.mesh.backend.ctx
        h = C.c_void_p()
        L.check(L.lib().bh_euler_create(sc.handle, int(degree), float(gamma), per.shape[0],
                                        per.ctypes.data if per.size else None,
                                        f.maps.ctypes.data, f.ndofs, C.byref(h)),
                self.ctx, "bh_euler_create")
        self.handle = h

    def residual(self, u: np.ndarray, b: np.ndarray | None = None, terms: int = 3,
                 face_coef: float = -1.0) -> np.ndarray:
        """b += ∫flux_Ω (terms & 1) + face_coef Σ∫flux_Γ (terms & 2); host arrays."""
        u = np.ascontiguousarray(u, dtype=np.float64)
        if u.size != self.ndofs:
            raise ValueError("u has the wrong size")
        if b is None:
            b = np.zeros(self.ndofs, dtype=np.float64)
        if b.dtype != np.float64 or not b.flags.c_contiguous or b.size != self.ndofs:
            raise ValueError("b must be a contiguous float64 vector of ndofs entries")
        L.check(L.lib().bh_euler_residual(self.handle, terms, float(face_coef), u.ctypes.data,
                                          b.ctypes.data, L.HOST), self.ctx, "bh_euler_residual")
        return b

    def residual_device(self, u_ptr: int, b_ptr: int, terms: int = 3, face_coef: float = -1.0):
        L.check(L.lib().bh_euler_residual(self.handle, terms, float(face_coef), u_ptr, b_ptr,
                                          L.DEVICE), self.ctx, "bh_euler_residual")

    def step(self, u: np.ndarray, dt: float) -> np.ndarray:
        """forward_euler(u, compute_residual, t, Δt) (covo.jl:14-33, 152)."""
        u = np.ascontiguousarray(u, dtype=np.float64)
        if u.size != self.ndofs:
            raise ValueError("u has the wrong size")
        out = np.empty_like(u)
        L.check(L.lib().bh_euler_step(self.handle, float(dt), u.ctypes.data, out.ctypes.data,
                                      L.HOST), self.ctx, "bh_euler_step")
        return out

    def step_device(self, u_ptr: int, out_ptr: int, dt: float):
        L.check(L.lib().bh_euler_step(self.handle, float(dt), u_ptr, out_ptr, L.DEVICE), self.ctx,
                "bh_euler_step")

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            L.lib().bh_euler_destroy(h)
            self.handle = None


def _sign_and_core(expr):
    fs = _factors(expr, [])
    sign = 1.0
    core = [f for f in fs if not isinstance(f, Const)]
    for f in fs:
        if isinstance(f, Const):
            sign *= f.v
    return sign, core


def _is_tests(t, wrap):
    return (isinstance(t, tuple) and len(t) == 4 and all(
        isinstance(x, wrap) and isinstance(x.a, Placeholder) and x.a.role == "test"
        and x.a.field == i for i, x in enumerate(t)))


def lower_euler_dg(items, V=None):
    """Recognise the COVO linear form.  Returns (gamma, u, degree, periodic
    domains, terms bitmask, face coefficient) or None when the form is not it.
    With the test space V given, the state's space must be numbered as V and
    every measure must lie on V's mesh (the kernel reads u in V's numbering)."""
    gamma = u = degree = fsign = None
    terms = 0
    periodic = []
    for it in items:
        sign, core = _sign_and_core(it.expr)
        if len(core) != 1 or not isinstance(core[0], Compose):
            return None
        cp = core[0]
        dom = it.measure.domain
        if isinstance(cp.f, EulerFlux):
            if not isinstance(dom, CellDomain) or sign != 1.0 or len(cp.args) != 2:
                return None
            uu, grads = cp.args
            if not isinstance(uu, FEFunction) or not _is_tests(grads, Grad):
                return None
            if terms & 1:
                raise NotImplementedError("the volume integral appears twice")
            terms |= 1
        elif isinstance(cp.f, RoeFlux):
            if sign not in (1.0, -1.0) or len(cp.args) != 4:
                return None
            if fsign is not None and sign != fsign:
                raise NotImplementedError("face integrals with different signs")
            fsign = sign
            a0, a1, a2, a3 = cp.args
            if not (isinstance(a0, Side) and a0.which == "n" and isinstance(a0.a, FEFunction)
                    and isinstance(a1, Side) and a1.which == "p" and a1.a is a0.a
                    and _is_tests(a2, Jump) and isinstance(a3, Side) and a3.which == "n"
                    and isinstance(a3.a, FaceNormal)):
                return None
            uu = a0.a
            if isinstance(dom, InteriorFaceDomain):
                if terms & 2:
                    raise NotImplementedError("the interior-face integral appears twice")
                terms |= 2
            elif isinstance(dom, BoundaryFaceDomain) and dom.bc is not None:
                if any(d is dom for d in periodic):
                    raise NotImplementedError("a periodic face integral appears twice")
                periodic.append(dom)
            else:
                return None
        else:
            return None
        if u is not None and uu is not u:
            raise NotImplementedError("one FEFunction across the integrals")
        if gamma is not None and cp.f.gamma != gamma:
            raise NotImplementedError("one γ across the integrals")
        if degree is not None and it.measure.degree != degree:
            raise NotImplementedError("one quadrature degree across the integrals")
        u, gamma, degree = uu, cp.f.gamma, it.measure.degree
    if u is None:
        return None
    if V is not None:
        from .assembler import Fields
        fv, fu = Fields(V), Fields(u.space)
        if fu.sizes != fv.sizes or not np.array_equal(fu.maps, fv.maps):
            raise NotImplementedError("the state's space must be numbered as the test space")
        for it in items:
            if it.measure.domain.mesh is not fv.scalar.mesh:
                raise ValueError("a measure's mesh differs from the space's mesh")
    if periodic and not terms & 2:
        raise NotImplementedError("periodic faces without the interior faces")
    return gamma, u, degree, periodic, terms, (fsign if fsign is not None else -1.0)


def euler_plan(V, gamma, degree, periodic):
    cache = V.__dict__.setdefault("_euler_plans", {})
    key = (float(gamma), int(degree), tuple(id(d) for d in periodic))
    if key not in cache:
        cache[key] = (EulerDG(V, gamma, degree, periodic), list(periodic))
    return cache[key][0]
