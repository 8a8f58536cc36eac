"""L2 projection onto a continuous space (src/feoperator/projection_newapi.jl:26-61).

projection_l2!(u, f, dΩ) solves M x = b with M = ∫(u v)dΩ from
assemble_bilinear and b = ∫(f v)dΩ from assemble_linear!, both on the GPU, and
stores x in u.  The sparse solve stays on the host, as the reference's
`__solve!` for its CPU backend (a direct solve of the assembled CSC matrix).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse.linalg as spla

from .assembler import assemble_bilinear, assemble_linear_
from .fespace import TestFESpace
from .forms import CellDomain, Measure, integrate


def projection_l2_(u, f, where, degree: int | None = None, *, mass=None) -> None:
    """projection_l2!(u, f, dΩ; mass), projection_l2!(u, f, mesh[, degree])
    (projection_newapi.jl:26-61): the default degree for a mesh is
    2 * (degree of u's function space) + 1.  `mass` replaces the assembled
    mass matrix (a scipy sparse matrix or a callable solving M x = b)."""
    U = u.space
    if isinstance(where, Measure):
        dΩ = where
    else:
        if degree is None:
            degree = 2 * U.fs.degree + 1
        dΩ = Measure(CellDomain(where), degree)
    V = TestFESpace(U)
    b = np.zeros(U.get_ndofs(), dtype=np.float64)
    assemble_linear_(b, lambda v: integrate(f * v, dΩ), V)
    if mass is None:
        A = assemble_bilinear(lambda a, v: integrate(a * v, dΩ), U, V)
        x = spla.spsolve(A.tocsc(), b)
    elif callable(mass):
        x = mass(b)
    else:
        x = spla.spsolve(mass.tocsc(), b)
    u.set_dof_values_(x)
