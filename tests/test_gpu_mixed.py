"""GPU parity of triangle and mixed triangle/quad assembly (mass + Laplacian)
against the oracle: FP64 within 1e-12 * max|A|, FP32 matrix cores within
1e-6 * max|A|, pattern bit-exact.  Covers P1..P4 triangles (the P4 basis and
the triangle rules are pinned by the oracle's own tests) on meshes mixing both
cell types, one kernel launch per cell type plus one row gather."""
import numpy as np
import pytest

import bcubehip as B
from oracle import oracle as O

pytestmark = pytest.mark.gpu
RTOL = 1e-12
RTOL_FP32 = 1e-6


def _om(m):
    return O.OMesh(m.dim, m.coords, m.celltypes, m.c2n_ptr, m.c2n)


def _plan(m, p, deg, kappa=None, mode="fp64"):
    U = B.TrialFESpace(B.FunctionSpace("Lagrange", p), m)
    dΩ = B.Measure(B.CellDomain(m), deg)
    if kappa is None:
        a = lambda u, v: B.integrate(B.dot(B.grad(u), B.grad(v)), dΩ) + B.integrate(u * v, dΩ)
    else:
        kc = B.MeshCellData(kappa)
        a = lambda u, v: (B.integrate(kc * B.dot(B.grad(u), B.grad(v)), dΩ)
                          + B.integrate(u * v, dΩ))
    return B.BilinearPlan(a, U, B.TestFESpace(U), mode)


def _ref(m, p, deg, kappa=None):
    return O.assemble_bilinear(_om(m), p, ["stiffness", "mass"], eta=kappa, degree=deg)


def _check(plan, ref, rtol, mode="fp64"):
    colptr, rowval, nzval, n = ref
    c, r = plan.csc_structure()
    assert np.array_equal(c + 1, colptr) and np.array_equal(r + 1, rowval)
    vals = plan.assemble_host(mode=mode) if mode != "fp64" else plan.assemble_host()
    err = np.abs(vals - nzval).max() / np.abs(nzval).max()
    assert err <= rtol, err
    return err


@pytest.mark.parametrize("p,deg", [(1, 4), (2, 4), (2, 5), (3, 6), (4, 8)])
def test_triangle_mesh_vs_oracle(p, deg):
    m, _ = B.mixed_tri_quad_mesh(8, block=8, p_split=1.0, seed=0)
    B.jitter_nodes(m, 0.2, 1.0 / 8, seed=p)
    assert np.all(m.celltypes == 2)
    kappa = np.random.default_rng(p).uniform(0.5, 2.0, m.ncells)
    plan = _plan(m, p, deg, kappa)
    print(f"Tri P{p} deg {deg}: rel err {_check(plan, _ref(m, p, deg, kappa), RTOL):.2e}")


@pytest.mark.parametrize("p,deg", [(1, 1), (1, 2), (1, 3), (2, 2), (2, 3)])
def test_triangle_low_degree_rules_fp64_and_fp32(p, deg):
    """Measure degrees 1-3 on triangles (FEMQuad's 1-, 3- and 4-point rules,
    one with a negative weight) through the FP64 and FP32 element kernels."""
    m, _ = B.mixed_tri_quad_mesh(8, block=8, p_split=1.0, seed=1)
    B.jitter_nodes(m, 0.2, 1.0 / 8, seed=p + 10)
    kappa = np.random.default_rng(p).uniform(0.5, 2.0, m.ncells)
    plan = _plan(m, p, deg, kappa)
    ref = _ref(m, p, deg, kappa)
    e64 = _check(plan, ref, RTOL)
    e32 = _check(plan, ref, RTOL_FP32, mode="fp32")
    print(f"Tri P{p} deg {deg}: FP64 {e64:.2e}, FP32 {e32:.2e}")


@pytest.mark.parametrize("p,deg", [(1, 2), (1, 4), (2, 3), (2, 4), (3, 6), (4, 8)])
def test_mixed_mesh_vs_oracle_fp64_and_fp32(p, deg):
    m, _ = B.mixed_tri_quad_mesh(12, block=3, p_split=0.5, seed=p)
    B.jitter_nodes(m, 0.2, 1.0 / 12, seed=p + 5)
    kappa = np.exp(np.random.default_rng(p).normal(0.0, 0.5, m.ncells))
    plan = _plan(m, p, deg, kappa)
    ref = _ref(m, p, deg, kappa)
    e64 = _check(plan, ref, RTOL)
    e32 = _check(plan, ref, RTOL_FP32, mode="fp32")
    print(f"mixed P{p}/Q{p} deg {deg}: FP64 {e64:.2e}, FP32 {e32:.2e}")


def test_mixed_determinism_and_symmetry():
    m, _ = B.mixed_tri_quad_mesh(16, block=4, p_split=0.5, seed=9)
    plan = _plan(m, 3, 6)
    a = plan.assemble_host()
    b = plan.assemble_host()
    assert np.array_equal(a, b)
    A = plan.to_csc(a)
    assert abs(A - A.T).max() <= 1e-14 * abs(A).max()


# Every (cell type, order, rule) of the FP32 GEMM element kernel, on a uniform
# mesh (cells in mesh order) and, for quads and triangles, inside a mixed mesh
# (cells picked through their global ids, rows at the larger element's
# stride) where both cell types have a kernel for the order and degree.
@pytest.mark.parametrize("shape,p,deg", [
    ("quad", 1, 2), ("quad", 1, 4), ("quad", 2, 2), ("quad", 2, 4), ("quad", 3, 4),
    ("quad", 3, 6), ("quad", 4, 8),
    ("tri", 1, 4), ("tri", 2, 4), ("tri", 2, 5), ("tri", 3, 6), ("tri", 4, 8),
    ("hexa", 1, 2), ("hexa", 2, 2), ("hexa", 2, 5),
])
def test_gemm_element_kernel_every_rule_fp32(shape, p, deg):
    rng = np.random.default_rng(100 + p * 10 + deg)
    if shape == "hexa":
        m = B.hexa_mesh(6, 5, 4)
        B.jitter_nodes(m, 0.15, 0.2, seed=deg)
    elif shape == "quad":
        m = B.rectangle_mesh(11, 9)
        B.jitter_nodes(m, 0.2, 0.1, seed=deg)
    else:
        m, _ = B.mixed_tri_quad_mesh(8, block=8, p_split=1.0, seed=p)
        B.jitter_nodes(m, 0.2, 1.0 / 8, seed=deg)
    kappa = rng.uniform(0.5, 2.0, m.ncells)
    plan = _plan(m, p, deg, kappa, mode="fp32")
    e32 = _check(plan, _ref(m, p, deg, kappa), RTOL_FP32, mode="fp32")
    msg = f"{shape} order {p} degree {deg}: uniform FP32 {e32:.2e}"
    # mixed meshes where both cell types have a kernel for (p, degree)
    if shape != "hexa" and (p, deg) in {(1, 4), (2, 4), (2, 5), (3, 6), (4, 8)}:
        mm, _ = B.mixed_tri_quad_mesh(10, block=2, p_split=0.5, seed=deg)
        B.jitter_nodes(mm, 0.2, 0.1, seed=p)
        k2 = rng.uniform(0.5, 2.0, mm.ncells)
        e2 = _check(_plan(mm, p, deg, k2, mode="fp32"), _ref(mm, p, deg, k2), RTOL_FP32, mode="fp32")
        msg += f", mixed FP32 {e2:.2e}"
    print(msg)


@pytest.mark.parametrize("p,deg,src", [(1, 2, "one"), (1, 4, "x"), (2, 4, "sinsin"),
                                       (3, 6, "sinsin"), (4, 8, "x")])
def test_mixed_mesh_linear_form_vs_oracle(p, deg, src):
    """∫f v dΩ on a mixed mesh: one launch per colour and cell type, the
    PhysicalFunction sampled at points padded to the larger rule."""
    m, _ = B.mixed_tri_quad_mesh(12, block=3, p_split=0.5, seed=p + 20)
    B.jitter_nodes(m, 0.2, 1.0 / 12, seed=p + 21)
    U = B.TrialFESpace(B.FunctionSpace("Lagrange", p), m)
    dΩ = B.Measure(B.CellDomain(m), deg)
    f = {"one": 1.0,
         "x": B.PhysicalFunction(lambda x: x[:, 0]),
         "sinsin": B.PhysicalFunction(lambda x: np.sin(np.pi * x[:, 0]) * np.sin(np.pi * x[:, 1]))}
    b = B.assemble_linear(lambda v: B.integrate(f[src] * v, dΩ), B.TestFESpace(U))
    ref = O.assemble_linear(_om(m), p, src, degree=deg)
    err = np.abs(b - ref).max() / np.abs(ref).max()
    print(f"mixed P{p} deg {deg} f={src}: rel err {err:.2e}")
    assert err <= 1e-12


def test_c3_reduced_mesh_every_bucket_vs_oracle():
    """The C3 workload reduced to 96^2 quads: 16x16 blocks split into
    triangles with probability 1/3 (seed 0), block orders p in {1,2,3,4} with
    probabilities {0.4,0.3,0.2,0.1} (seed 1); each order bucket (continuous
    P_p/Q_p on the union of its blocks, ∫∇u⋅∇v + ∫uv, degree max(2p, 4))
    against the oracle, FP64 and FP32."""
    n = 96
    m, blocks = B.mixed_tri_quad_mesh(n, block=16, p_split=1.0 / 3.0, seed=0)
    nb = int(blocks.max()) + 1
    border = np.random.default_rng(1).choice([1, 2, 3, 4], size=nb, p=[0.4, 0.3, 0.2, 0.1])
    order = border[blocks]
    assert (m.celltypes == 2).any() and (m.celltypes == 0).any()
    seen = 0
    for p in (1, 2, 3, 4):
        cells = np.flatnonzero(order == p)
        if cells.size == 0:
            continue
        seen += 1
        sub = B.sub_mesh(m, cells)
        deg = max(2 * p, 4)
        plan = _plan(sub, p, deg)
        ref = _ref(sub, p, deg)
        e64 = _check(plan, ref, RTOL)
        e32 = _check(plan, ref, RTOL_FP32, mode="fp32")
        print(f"C3 96^2 bucket P{p}: {cells.size} cells, fp64 {e64:.2e}, fp32 {e32:.2e}")
    assert seen >= 3
