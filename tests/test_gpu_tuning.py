"""Launch shapes: every candidate shape of the FP32 GEMM element kernel and of
the row gather gives bitwise the same CSR values as the default, and the
autotuner's choice does too (Q2 hexa, mixed P2 triangles/quads, Q1 quads)."""
import numpy as np
import pytest

import bcubehip as B

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
    else:
        m = B.rectangle_mesh(41, 41)
        B.jitter_nodes(m, 0.2, 1.0 / 40, seed=7)
        p, deg = 1, 2
    kappa = np.random.default_rng(3).uniform(0.5, 2.0, m.ncells)
    U = B.TrialFESpace(B.FunctionSpace("Lagrange", p), m)
    dΩ = B.Measure(B.CellDomain(m), deg)
    kc = B.MeshCellData(kappa)
    form = lambda u, v: B.integrate(kc * B.dot(B.grad(u), B.grad(v)), dΩ) + B.integrate(u * v, dΩ)
    return m, U, form


@pytest.mark.parametrize("name", ["hexa_q2", "mixed_p2", "quad_q1"])
@pytest.mark.parametrize("mode", ["fp32", "fp64"])
def test_every_shape_bitwise_equal(name, mode):
    m, U, form = _case(name)
    plan = B.BilinearPlan(form, U, B.TestFESpace(U), mode)
    default = plan.assemble_host().copy()
    # the element + row-gather shapes (FP64 Q2 on a grid defaults to the
    # one-pass kernel, which sums each entry in another fixed order)
    plan.set_tuning(8, 0, 0)
    ref = plan.assemble_host().copy()
    assert np.abs(default - ref).max() <= 1e-13 * np.abs(ref).max()
    tried = 0
    for waves in (8, 12):
        for lanes in (0, 4, 8, 16, 32):
            for rows in ((0,) if lanes == 0 else (1, 2)):
                try:
                    plan.set_tuning(waves, lanes, rows)
                except B.BcubeHipError:
                    continue
                vals = plan.assemble_host()
                assert np.array_equal(vals, ref), (name, mode, waves, lanes, rows)
                tried += 1
    print(name, mode, "shapes tried", tried)
    assert tried >= 4


@pytest.mark.parametrize("name", ["hexa_q2", "mixed_p2"])
def test_autotuned_shapes_bitwise_equal(name):
    m, U, form = _case(name)
    ref = B.BilinearPlan(form, U, B.TestFESpace(U), "fp32").assemble_host().copy()
    m.backend.set_autotune(True)
    try:
        plan = B.BilinearPlan(form, U, B.TestFESpace(U), "fp32")
        assert plan.tuning() is None
        vals = plan.assemble_host()
        t = plan.tuning()
        print(name, "tuned", t)
        assert t is not None and t["gather_lanes"] in (4, 8, 16, 32) and t["elem_waves"] in (8, 12)
        assert t["elem_ms"] > 0 and t["gather_ms"] > 0
        assert np.array_equal(vals, ref)
        assert np.array_equal(plan.assemble_host(), ref)
    finally:
        m.backend.set_autotune(False)
