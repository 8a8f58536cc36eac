"""Upwind DG transport residual on the HIP backend, and the explicit time loop
that re-assembles it every stage.

The residual is the linear form (with Ω = CellDomain, Γ = InteriorFaceDomain)

    l(v) = ∫((c*u)⋅∇(v))dΩ - ∫(flux ∘ (side⁻(u), side⁺(u), jump(v), side⁻(nΓ)))dΓ
    flux(ui, uj, δv, n) = (c⋅n > 0 ? (c⋅n)*ui : (c⋅n)*uj) * δv

that assemble_linear(l, V) (src/assembler/assembler.jl:448-533, faces
:768-800) evaluates cell by cell and face by face; here bh_dg_transport_residual
computes it for several fields at once on the GPU.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L


def _torch():
    try:
        import torch
        return torch
    except ImportError:
        return None


class DGTransport:
    """Plan of the DG transport residual on a discontinuous space V.

    velocity: the constant c (3 values); degree: Measure degree of both Ω and Γ
    (order+1 Gauss points per axis); mode "fp64" or "fp32" (FP32 fields and
    arithmetic, geometry computed in FP64).  Fields are laid out field-major:
    a (nfields, ndofs) array, each row one field's dof vector.
    """

    def __init__(self, V, velocity, degree: int, mode: str = "fp64"):
        V = getattr(V, "U", V)
        if V.continuous:
            raise ValueError("DGTransport needs a discontinuous space")
        if mode not in ("fp64", "fp32"):
            raise ValueError("mode must be 'fp64' or 'fp32'")
        self.V = V
        self.mode = mode
        self.degree = int(degree)
        self.velocity = np.ascontiguousarray(velocity, dtype=np.float64).reshape(3)
        self.dtype = np.float32 if mode == "fp32" else np.float64
        self._ctx = V.mesh.backend.ctx
        h = C.c_void_p()
        L.check(L.lib().bh_dg_create(V.handle, self.degree, self.velocity.ctypes.data,
                                     L.MODE[mode], C.byref(h)), self._ctx, "bh_dg_create")
        self._h = h
        nd, nf, nl = C.c_int64(), C.c_int64(), C.c_int32()
        L.check(L.lib().bh_dg_info(h, C.byref(nd), C.byref(nf), C.byref(nl)), self._ctx,
                "bh_dg_info")
        self.ndofs, self.ninterior_faces, self.ndofs_per_cell = nd.value, nf.value, nl.value

    @property
    def handle(self):
        return self._h

    def update_geometry(self):
        """Recompute the geometric factors after Mesh.set_nodes."""
        L.check(L.lib().bh_dg_update_geometry(self._h), self._ctx, "bh_dg_update_geometry")

    def _torch_dtype(self):
        t = _torch()
        return t.float32 if self.mode == "fp32" else t.float64

    def residual(self, u):
        """r = l(v) for u of shape (ndofs,) or (nfields, ndofs).

        A numpy array is copied to the device and back; a torch tensor on the
        GPU stays there (result is a new tensor on the same device)."""
        t = _torch()
        if t is not None and isinstance(u, t.Tensor):
            if not u.is_cuda or u.dtype != self._torch_dtype() or not u.is_contiguous():
                raise ValueError("u must be a contiguous GPU tensor of the plan's dtype")
            if u.shape[-1] != self.ndofs:
                raise ValueError(f"u has {u.shape[-1]} dofs per field, the space has {self.ndofs}")
            r = t.empty_like(u)
            nf = 1 if u.dim() == 1 else u.shape[0]
            self.residual_device(nf, u.data_ptr(), r.data_ptr())
            return r
        u = np.ascontiguousarray(u, dtype=self.dtype)
        if u.shape[-1] != self.ndofs:
            raise ValueError(f"u has {u.shape[-1]} dofs per field, the space has {self.ndofs}")
        r = np.empty_like(u)
        nf = 1 if u.ndim == 1 else u.shape[0]
        L.check(L.lib().bh_dg_transport_residual(self._h, nf, u.ctypes.data, r.ctypes.data,
                                                 L.HOST), self._ctx, "bh_dg_transport_residual")
        return r

    def residual_device(self, nfields: int, u_ptr: int, r_ptr: int):
        L.check(L.lib().bh_dg_transport_residual(self._h, nfields, u_ptr, r_ptr, L.DEVICE),
                self._ctx, "bh_dg_transport_residual")

    def minv_axpy(self, nfields: int, r_ptr: int, x1_ptr: int, a1: float, out1_ptr: int,
                  x2_ptr: int | None = None, a2: float = 0.0, out2_ptr: int | None = None):
        """out1 = x1 + a1 M⁻¹r (and out2 = x2 + a2 M⁻¹r), device pointers."""
        L.check(L.lib().bh_dg_minv_axpy(self._h, nfields, r_ptr, x1_ptr, float(a1), out1_ptr,
                                        x2_ptr, float(a2), out2_ptr), self._ctx,
                "bh_dg_minv_axpy")

    def stage(self, nfields: int, u_ptr: int, x1_ptr: int, a1: float, out1_ptr: int,
              x2_ptr: int | None = None, a2: float = 0.0, out2_ptr: int | None = None,
              work_ptr: int | None = None):
        """out1 = x1 + a1 M⁻¹ l(u) (and out2 = x2 + a2 M⁻¹ l(u)) in one pass
        (bh_dg_transport_stage); u must not alias an output."""
        L.check(L.lib().bh_dg_transport_stage(self._h, nfields, u_ptr, x1_ptr, float(a1),
                                              out1_ptr, x2_ptr, float(a2), out2_ptr, work_ptr),
                self._ctx, "bh_dg_transport_stage")

    def _rk4_args(self, u, work):
        t = _torch()
        if t is None or not isinstance(u, t.Tensor) or not u.is_cuda:
            raise ValueError("rk4_step works on a GPU tensor")
        if work is None:
            work = tuple(t.empty_like(u) for _ in range(3))
        nf = 1 if u.dim() == 1 else u.shape[0]
        return work, nf

    def rk4_step(self, u, dt: float, work=None):
        """One classical RK4 step of M du/dt = l(u) in place on a GPU tensor u
        (nfields, ndofs): four stages, each a residual assembly and its M⁻¹
        updates in one pass (bh_dg_transport_stage), the stage inputs in two
        alternating buffers.  `work` = three tensors shaped like u (allocated
        if None) -- pass them to keep the call allocation-free (and
        graph-capturable).  Bitwise equal to rk4_step_two_pass."""
        work, nf = self._rk4_args(u, work)
        r, acc, us = work
        U, S2, A, S1 = u.data_ptr(), r.data_ptr(), acc.data_ptr(), us.data_ptr()
        # the work array of the two-launch form (FP64 Q3) is each stage's
        # second output, or the free stage buffer in the last stage
        self.stage(nf, U, U, dt / 6, A, U, dt / 2, S1, S1)
        self.stage(nf, S1, A, dt / 3, A, U, dt / 2, S2, S2)
        self.stage(nf, S2, A, dt / 3, A, U, dt, S1, S1)
        self.stage(nf, S1, A, dt / 6, U, None, 0.0, None, S2)
        return work

    def rk4_step_two_pass(self, u, dt: float, work=None):
        """The same RK4 step as four residual assemblies and four M⁻¹ updates."""
        work, nf = self._rk4_args(u, work)
        r, acc, us = work
        U, R, A, S = u.data_ptr(), r.data_ptr(), acc.data_ptr(), us.data_ptr()
        self.residual_device(nf, U, R)
        self.minv_axpy(nf, R, U, dt / 6, A, U, dt / 2, S)
        self.residual_device(nf, S, R)
        self.minv_axpy(nf, R, A, dt / 3, A, U, dt / 2, S)
        self.residual_device(nf, S, R)
        self.minv_axpy(nf, R, A, dt / 3, A, U, dt, S)
        self.residual_device(nf, S, R)
        self.minv_axpy(nf, R, A, dt / 6, U)
        return work

    def __del__(self):
        try:
            if getattr(self, "_h", None) is not None:
                L.lib().bh_dg_destroy(self._h)
        except Exception:
            pass


def dg_dof_coordinates(V) -> np.ndarray:
    """Reference-node positions of a discontinuous Q_p space mapped to physical
    space, (ndofs, 3), for building nodal initial conditions: node i of cell c
    is x_c(ξ_i) with ξ_i on the uniform tensor grid, first index fastest."""
    V = getattr(V, "U", V)
    mesh = V.mesh
    p = V.fs.degree
    N = p + 1
    xi1 = -1.0 + 2.0 * np.arange(N) / p
    pts = np.array([[xi1[i1], xi1[i2], xi1[i3]]
                    for i3 in range(N) for i2 in range(N) for i1 in range(N)])
    X = mesh.nodes()
    from . import mesh as M
    ct = int(mesh.celltypes[0])
    if ct == L.CELLTYPE["Hexa8"]:
        ref = M._H8_VERTS.astype(float)
        w = np.prod((1 + pts[:, None, :] * ref[None]) / 2, axis=2)          # (nl, 8)
    elif ct == L.CELLTYPE["Hexa27"]:
        ref = M._H27_REF.astype(float)

        def l1(t, a):
            return np.where(a < 0, 0.5 * t * (t - 1), np.where(a > 0, 0.5 * t * (t + 1), 1 - t * t))
        w = np.prod(l1(pts[:, None, :], ref[None]), axis=2)                  # (nl, 27)
    else:
        raise NotImplementedError("dg_dof_coordinates: Hexa8/Hexa27 meshes only")
    c2n = mesh.c2n.reshape(mesh.ncells, -1)
    return np.einsum("qk,ckd->cqd", w, X[c2n]).reshape(-1, 3)
