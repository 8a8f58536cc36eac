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


@pytest.mark.parametrize("name", ["mixed_p2", "quad_q1"])
def test_tuning_cache_persists_choice(name, tmp_path):
    """The autotuner's choice is written to the cache file and taken from it by
    a new context without timing (bh_ctx_set_tuning_cache); values bitwise
    equal."""
    path = str(tmp_path / "tune.txt")
    m, U, form = _case(name)
    be = m.backend
    be.set_autotune(True)
    be.set_tuning_cache(path)
    try:
        plan = B.BilinearPlan(form, U, B.TestFESpace(U), "fp32")
        v1 = plan.assemble_host().copy()
        t1 = plan.tuning()
    finally:  # the default backend is shared with the other tests
        be.set_autotune(False)
        be.set_tuning_cache(None)
    lines = [l.split() for l in open(path) if not l.startswith("#")]
    assert len(lines) == 1 and [int(x) for x in lines[0][1:4]] == [
        t1["elem_waves"], t1["gather_lanes"], t1["gather_rows"]]
    # a fresh context and mesh reading the same file: the cached shape, its
    # recorded times, no new entry
    be2 = B.BcubeBackendHIP()
    be2.set_autotune(True)
    be2.set_tuning_cache(path)
    m2, _, _ = _case(name)
    m2 = B.Mesh(m2.nodes(), m2.celltypes, [m2.cell_nodes(c) for c in range(m2.ncells)], be2)
    kappa = np.random.default_rng(3).uniform(0.5, 2.0, m2.ncells)
    p, deg = (2, 4) if name == "mixed_p2" else (1, 2)
    U2 = B.TrialFESpace(B.FunctionSpace("Lagrange", p), m2)
    dΩ = B.Measure(B.CellDomain(m2), deg)
    kc = B.MeshCellData(kappa)
    plan2 = B.BilinearPlan(lambda u, v: B.integrate(kc * B.dot(B.grad(u), B.grad(v)), dΩ)
                           + B.integrate(u * v, dΩ), U2, B.TestFESpace(U2), "fp32")
    v2 = plan2.assemble_host()
    t2 = plan2.tuning()
    assert {k: t2[k] for k in ("elem_waves", "gather_lanes", "gather_rows")} == {
        k: t1[k] for k in ("elem_waves", "gather_lanes", "gather_rows")}
    assert abs(t2["elem_ms"] - t1["elem_ms"]) <= 1e-3 * max(1.0, t1["elem_ms"])
    assert np.array_equal(v1, v2)
    assert len([l for l in open(path) if not l.startswith("#")]) == 1
    be2.set_tuning_cache(None)
