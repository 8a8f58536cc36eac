"""Offset templates of the row gather (bh_pattern_set_offset_templates): the
distinct element-row offset vectors in an LDS table and a template id beside
each pair id give bitwise the same CSR values as the per-entry offsets, on
structured Q2 hexahedra (122 templates), Q1/Q3 quads and a mixed P2 mesh, in
FP64 and FP32; FP64 values within 1e-12 of the oracle."""
import numpy as np
import pytest

import bcubehip as B
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _case(name):
    if name == "hexa_q2":
        m = B.hexa_mesh(11, 11, 11)
        B.jitter_nodes(m, 0.15, 0.1, seed=4)
        p, deg = 2, 5
    elif name == "mixed_p2":
        m, _ = B.mixed_tri_quad_mesh(24, block=4, p_split=0.5, seed=5)
        B.jitter_nodes(m, 0.2, 1.0 / 24, seed=6)
        p, deg = 2, 4
    elif name == "quad_q3":
        m = B.rectangle_mesh(21, 21)
        B.jitter_nodes(m, 0.2, 0.05, seed=8)
        p, deg = 3, 6
    else:
        m = B.rectangle_mesh(41, 41)
        B.jitter_nodes(m, 0.2, 1.0 / 40, seed=7)
        p, deg = 1, 2
    kappa = np.random.default_rng(3).uniform(0.5, 2.0, m.ncells)
    U = B.TrialFESpace(B.FunctionSpace("Lagrange", p), m)
    dΩ = B.Measure(B.CellDomain(m), deg)
    kc = B.MeshCellData(kappa)
    form = lambda u, v: B.integrate(kc * B.dot(B.grad(u), B.grad(v)), dΩ) + B.integrate(u * v, dΩ)
    return m, U, form, kappa, p, deg


@pytest.mark.parametrize("name", ["hexa_q2", "quad_q1", "quad_q3", "mixed_p2"])
@pytest.mark.parametrize("mode", ["fp64", "fp32"])
def test_templates_bitwise_equal(name, mode):
    m, U, form, kappa, p, deg = _case(name)
    plan = B.BilinearPlan(form, U, B.TestFESpace(U), mode)
    plan.set_tuning(8, 0, 0)  # element kernel + row gather (not the one-pass FP64 kernel)
    a = plan.assemble_host().copy()
    ntpl = plan.offset_templates()
    plan.set_offset_templates(False)
    assert plan.offset_templates() == 0
    b = plan.assemble_host().copy()
    plan.set_offset_templates(True)
    c = plan.assemble_host()
    print(f"{name} {mode}: {ntpl} templates", flush=True)
    if name == "hexa_q2":
        assert ntpl == 122
    if name in ("quad_q1", "quad_q3"):
        assert 0 < ntpl <= 255
    assert np.array_equal(a, b) and np.array_equal(a, c)
    if mode == "fp64":
        o = O.OMesh(m.dim, m.coords, m.celltypes, m.c2n_ptr, m.c2n)
        _, _, nz, _ = O.assemble_bilinear(o, p, ["stiffness", "mass"], eta=kappa, degree=deg)
        assert np.abs(a - nz).max() <= 1e-12 * np.abs(nz).max()
