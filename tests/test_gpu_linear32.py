"""GPU parity of the FP32 linear forms (element vectors b_c = Φᵀ g_c on the
matrix cores, FP64 geometry and coefficients rounded once to FP32, summed into
b per dof in FP64 in increasing cell order) against the oracle's FP64
assemble_linear (assembler.jl:523-533, 855-865).

Tolerance: max|b - b_oracle| <= 1e-6 * max|b_oracle| (normwise over the
vector), the FP32 bound of the bilinear forms.  Repeated calls are bitwise
identical; b += l(v) keeps the caller's values.
"""
import numpy as np
import pytest

import bcubehip as B
from oracle import oracle as O

pytestmark = pytest.mark.gpu
RTOL_FP32 = 1e-6

SRC = {"one": lambda: 1.0,
       "x": lambda: B.PhysicalFunction(lambda x: x[:, 0]),
       "sinsin": lambda: B.PhysicalFunction(lambda x: np.sin(np.pi * x[:, 0]) * np.sin(np.pi * x[:, 1]))}


def _om(m):
    return O.OMesh(m.dim, m.coords, m.celltypes, m.c2n_ptr, m.c2n)


def _mesh(kind, n, seed):
    if kind == "quad":
        m = B.rectangle_mesh(n + 1, n + 1)
        B.jitter_nodes(m, 0.2, 1.0 / n, seed=seed)
    elif kind == "mixed":
        m, _ = B.mixed_tri_quad_mesh(n, block=3, p_split=0.5, seed=seed)
        B.jitter_nodes(m, 0.2, 1.0 / n, seed=seed + 1)
    elif kind == "tri":
        m, _ = B.mixed_tri_quad_mesh(n, block=n, p_split=1.0, seed=seed)
        B.jitter_nodes(m, 0.2, 1.0 / n, seed=seed + 1)
    elif kind == "hexa":
        m = B.hexa_mesh(n + 1, n + 1, n + 1)
        B.jitter_nodes(m, 0.15, 1.0 / n, seed=seed)
    elif kind == "quad9":
        m = B.rectangle_mesh(n + 1, n + 1, order=2)
        B.sinusoidal_map(m, 0.05, seed=seed)
    else:
        m = B.hexa27_mesh(n + 1, n + 1, n + 1)
        B.sinusoidal_map(m, 0.05, seed=seed)
    return m


def _linear(m, p, deg, src, mode, b0=None):
    V = B.TestFESpace(B.TrialFESpace(B.FunctionSpace("Lagrange", p), m))
    dΩ = B.Measure(B.CellDomain(m), deg)
    f = SRC[src]()
    b = np.zeros(V.get_ndofs()) if b0 is None else b0.copy()
    B.assemble_linear_(b, lambda v: B.integrate(f * v, dΩ), V, mode=mode)
    return b, V


CASES = [("quad", 12, 1, 2, "one"), ("quad", 12, 1, 3, "sinsin"), ("quad", 10, 2, 4, "x"),
         ("quad", 8, 3, 6, "sinsin"), ("quad", 8, 4, 8, "x"),
         ("tri", 8, 1, 2, "x"), ("tri", 8, 2, 4, "sinsin"), ("tri", 6, 3, 6, "x"),
         ("tri", 6, 4, 8, "sinsin"), ("mixed", 12, 1, 4, "x"), ("mixed", 12, 2, 4, "sinsin"),
         ("mixed", 12, 4, 8, "x"), ("hexa", 5, 1, 2, "x"), ("hexa", 5, 1, 4, "one"),
         ("hexa", 4, 2, 4, "sinsin"), ("hexa", 4, 2, 5, "x"), ("quad9", 6, 2, 4, "x"),
         ("quad9", 6, 1, 3, "sinsin"), ("hexa27", 3, 2, 5, "x"), ("hexa27", 3, 1, 3, "one")]


@pytest.mark.parametrize("kind,n,p,deg,src", CASES)
def test_fp32_linear_form_vs_oracle(kind, n, p, deg, src):
    print(f"case: fp32 linear {kind} n={n} P{p} degree {deg} f={src}", flush=True)
    m = _mesh(kind, n, seed=p + deg)
    b32, _ = _linear(m, p, deg, src, "fp32")
    ref = O.assemble_linear(_om(m), p, src, degree=deg)
    err = np.abs(b32 - ref).max() / np.abs(ref).max()
    b64, _ = _linear(m, p, deg, src, "fp64")
    e64 = np.abs(b64 - ref).max() / np.abs(ref).max()
    print(f"{kind} P{p} degree {deg} f={src}: fp32 {err:.2e}, fp64 {e64:.2e}")
    assert 0 < err <= RTOL_FP32
    assert e64 <= 1e-12


def test_fp32_linear_form_repeatable_and_accumulates():
    print("case: fp32 linear repeat / accumulate", flush=True)
    m = _mesh("hexa", 6, seed=4)
    b1, V = _linear(m, 2, 4, "sinsin", "fp32")
    b2, _ = _linear(m, 2, 4, "sinsin", "fp32")
    assert np.array_equal(b1, b2)
    b0 = np.random.default_rng(0).uniform(-1, 1, V.get_ndofs())
    b3, _ = _linear(m, 2, 4, "sinsin", "fp32", b0=b0)
    # b += l(v): the caller's value first, then the element vectors in cell order
    assert np.allclose(b3, b0 + b1, rtol=0, atol=1e-15 * np.abs(b1).max() * 8)


def test_fp32_linear_form_cell_coefficient_and_device_path():
    """A per-cell coefficient κ_c (MeshCellData) against the FP64 GPU result,
    and the device-resident call writing into a torch tensor equal to the
    host call bit for bit."""
    import torch
    from bcubehip import _lib as L
    print("case: fp32 linear cell coefficient / device", flush=True)
    m = _mesh("quad", 16, seed=7)
    kappa = np.random.default_rng(3).uniform(0.5, 2.0, m.ncells)
    V = B.TestFESpace(B.TrialFESpace(B.FunctionSpace("Lagrange", 3), m))
    dΩ = B.Measure(B.CellDomain(m), 6)
    kc = B.MeshCellData(kappa)
    l = lambda v: B.integrate(kc * v, dΩ)
    b64 = np.zeros(V.get_ndofs())
    B.assemble_linear_(b64, l, V, mode="fp64")
    b32 = np.zeros(V.get_ndofs())
    B.assemble_linear_(b32, l, V, mode="fp32")
    err = np.abs(b32 - b64).max() / np.abs(b64).max()
    print(f"κ_c v, Q3: fp32 vs fp64 {err:.2e}")
    assert 0 < err <= RTOL_FP32
    # device call: b and κ in HBM
    U = getattr(V, "U", V)
    bd = torch.zeros(V.get_ndofs(), dtype=torch.float64, device="cuda")
    kd = torch.from_numpy(kappa).cuda()
    term = (L.BhTerm * 1)(L.BhTerm(L.FORM_SOURCE, L.COEF_CELL, kd.data_ptr(), 1.0))
    L.check(L.lib().bh_assemble_linear(U.handle, 1, term, 6, L.MODE["fp32"], bd.data_ptr(),
                                       L.DEVICE), m.backend.ctx, "bh_assemble_linear")
    torch.cuda.synchronize()
    assert np.array_equal(bd.cpu().numpy(), b32)
