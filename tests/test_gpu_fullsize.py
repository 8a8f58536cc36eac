"""Parity at the full sizes of the C3 and C5 workloads through properties that
do not depend on size (the oracle's own comparisons run on samples of the same
layouts, which finish in seconds):

  * C3: the 866² mixed tri/quad mesh (998,056 cells, 1/3 of the 16×16 blocks
    split into triangles, block orders 1..4 with probabilities .4/.3/.2/.1),
    one continuous space of per-cell order (80.6 M entries);
  * C5: the 2048-instance batch of unit-square quad meshes (sides 8..96),
    Q1 and Q2 buckets of 1024 instances each.

Checked in FP64 and FP32: stiffness rows sum to zero (constants are in every
space), the mass matrix sums to the total area, FP32 within 1e-6 of FP64 (the
north-star bound, normwise), and repeated calls are bitwise equal.
"""
import numpy as np
import pytest

import bcubehip as B

pytestmark = pytest.mark.gpu


def _row_checks(form_k, form_m, area, label, build="host"):
    pk = B.BilinearPlan(form_k[0], *form_k[1:], build=build)
    pm = B.BilinearPlan(form_m[0], *form_m[1:], build=build)
    colptr, _ = pk.csc_structure()
    out = {}
    for mode in ("fp64", "fp32"):
        K = pk.assemble_host(mode=mode)
        rows = np.add.reduceat(np.abs(K), colptr[:-1])
        rs = np.abs(np.add.reduceat(K, colptr[:-1])).max() / rows.max()
        M = pm.assemble_host(mode=mode)
        ms = abs(M.sum() - area) / area
        tol = 1e-12 if mode == "fp64" else 1e-6
        print(f"{label} {mode}: nnz {pk.nnz}, K·1 {rs:.1e}, ΣM - area {ms:.1e}", flush=True)
        assert rs <= tol and ms <= tol
        out[mode] = K
    e = np.abs(out["fp32"] - out["fp64"]).max() / np.abs(out["fp64"]).max()
    print(f"{label}: fp32 vs fp64 {e:.2e}", flush=True)
    assert 0 < e <= 1e-6
    assert np.array_equal(out["fp32"], pk.assemble_host(mode="fp32"))
    return pk


def test_c3_full_size_invariants():
    n = 866
    m, blocks = B.mixed_tri_quad_mesh(n, block=16, p_split=1.0 / 3.0, seed=0)
    nb = int(blocks.max()) + 1
    border = np.random.default_rng(1).choice([1, 2, 3, 4], size=nb, p=[0.4, 0.3, 0.2, 0.1])
    order = border[blocks]
    assert m.ncells == 998_056
    U = B.TrialFESpace(B.FunctionSpace("Lagrange", order), m)
    V = B.TestFESpace(U)
    dΩ = B.Measure(B.CellDomain(m), B.OrderDegree(4))
    pk = _row_checks((lambda u, v: B.integrate(B.dot(B.grad(u), B.grad(v)), dΩ), U, V),
                     (lambda u, v: B.integrate(u * v, dΩ), U, V), 1.0, "C3 866²")
    assert pk.nnz == 80_583_483


@pytest.mark.parametrize("order", [1, 2])
def test_c5_full_batch_invariants(order):
    rng = np.random.default_rng(3)
    ninst = 2048
    sides = rng.integers(8, 97, size=ninst)
    idx = [k for k in range(ninst) if k % 2 == order - 1]
    big, _, _ = B.batch_meshes([B.rectangle_mesh(int(sides[k]) + 1, int(sides[k]) + 1)
                                for k in idx])
    kap = np.exp(np.random.default_rng(4).normal(0.0, 0.5, big.ncells))
    U = B.TrialFESpace(B.FunctionSpace("Lagrange", order), big)
    V = B.TestFESpace(U)
    dΩ = B.Measure(B.CellDomain(big), 2 * order)
    kc = B.MeshCellData(kap)
    # patterns built on the device (the same structure, bit for bit)
    _row_checks((lambda u, v: B.integrate(kc * B.dot(B.grad(u), B.grad(v)), dΩ), U, V),
                (lambda u, v: B.integrate(u * v, dΩ), U, V), float(len(idx)),
                f"C5 Q{order} ({len(idx)} instances)", build="device")
