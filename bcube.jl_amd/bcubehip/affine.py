"""AffineFESystem (src/assembler/affine_fe_system.jl:9-119): the system
a(u, v) = l(v) assembled on the GPU (assemble_bilinear, assemble_linear), with
the Dirichlet lifting of solve! -- d = the Dirichlet vector, b0 = b - A d,
homogeneous conditions on (A, b0), u = A⁻¹ b0 + d.  The solve is a host sparse
solve, as the reference's default `A \\ b`."""
from __future__ import annotations

import numpy as np
import scipy.sparse.linalg as spla

from .assembler import assemble_bilinear, assemble_linear
from .dirichlet import apply_homogeneous_dirichlet_, assemble_dirichlet_vector
from .forms import FEFunction


def default_linsolve_(y, A, x):
    """default_linsolve!(y, A, x) = y .= A \\ x (affine_fe_system.jl:79)."""
    y[:] = spla.spsolve(A.tocsc(), x)


class AffineFESystem:
    """AffineFESystem(a, l, U, V[, linsolve!]) (affine_fe_system.jl:45-76)."""

    def __init__(self, a, l, U, V, linsolve_=default_linsolve_):
        Us, Vs = getattr(U, "U", U), getattr(V, "U", V)
        if Us is not Vs:
            raise NotImplementedError("U and V must be defined on the same FESpace")
        self.A = assemble_bilinear(a, U, V)
        self.b = assemble_linear(l, V)
        self.U, self.V = U, V
        self.mesh = Us.mesh
        self.linsolve_ = linsolve_

    def solve_(self, u: FEFunction, t: float = 0.0) -> None:
        """solve!(u, system, t) (affine_fe_system.jl:100-119)."""
        Us = getattr(self.U, "U", self.U)
        A = self.A.copy()
        if getattr(Us, "dirichlet", None):
            from .assembler import pattern_of
            if A.nnz != pattern_of(Us).nnz:
                raise NotImplementedError("Dirichlet conditions on matrices with face terms")
            d = assemble_dirichlet_vector(self.U, self.V, self.mesh, t)
            b0 = self.b - A @ d
            apply_homogeneous_dirichlet_(A, b0, self.U, self.V, self.mesh)
        else:
            d = np.zeros(self.b.size)
            b0 = self.b.copy()
        x = np.zeros(self.b.size)
        self.linsolve_(x, A, b0)
        u.set_dof_values_(x + d)

    def solve(self, t: float = 0.0) -> FEFunction:
        """solve(system, t) (affine_fe_system.jl:87-98)."""
        u = FEFunction(self.U)
        self.solve_(u, t)
        return u
