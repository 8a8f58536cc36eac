"""The fused FP32 element + row-assembly kernel (one workgroup per pencil of
cells, rows complete within two steps summed in LDS, the rest by the row
gather over the spill rows) against the two-kernel path (element kernel + row
gather of every row): the same element values and the same summation order,
so the values must be bitwise equal; and against the oracle at 1e-6.  Meshes
with sizes that are not multiples of the pencil tiles leave partial pencils
and empty cell slots in the last step of a pencil."""
import numpy as np
import pytest

import bcubehip as B
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _om(m):
    return O.OMesh(m.dim, m.coords, m.celltypes, m.c2n_ptr, m.c2n)


def _hexa(nx, ny, nz, seed=0):
    m = B.hexa_mesh(nx + 1, ny + 1, nz + 1)
    B.jitter_nodes(m, 0.15, 1.0 / max(nx, ny, nz), seed=seed)
    return m


def _quad(n, seed=0):
    m = B.rectangle_mesh(n + 1, n + 1)
    B.jitter_nodes(m, 0.2, 1.0 / n, seed=seed)
    return m


def _plan(m, p, deg, mass=True, seed=1):
    eta = np.random.default_rng(seed).uniform(0.5, 2.0, m.ncells)
    U = B.TrialFESpace(B.FunctionSpace("Lagrange", p), m)
    dΩ = B.Measure(B.CellDomain(m), deg)
    ηc = B.MeshCellData(eta)
    if mass:
        a = lambda u, v: (B.integrate(ηc * B.dot(B.grad(u), B.grad(v)), dΩ)
                          + B.integrate(u * v, dΩ))
    else:
        a = lambda u, v: B.integrate(ηc * B.dot(B.grad(u), B.grad(v)), dΩ)
    return B.BilinearPlan(a, U, B.TestFESpace(U)), eta


def _fused_and_split(plan, eta):
    """FP32 values of the fused path and of the two-kernel path (device
    assembly, the fused one timed to show it ran)."""
    import torch
    be = plan.V.mesh.backend
    eta_d = torch.from_numpy(eta).cuda()
    out = {}
    for shape in (-1, 8, 0):
        plan.set_tuning(elem_waves=shape, mode="fp32")
        v = torch.full((plan.nnz,), float("nan"), dtype=torch.float64, device="cuda")
        be.set_timing(True)
        plan.assemble_device(v.data_ptr(), [eta_d.data_ptr(), None], "fp32")
        be.sync()
        _, kinds = be.timings_ex()
        be.set_timing(False)
        out[shape] = (v.cpu().numpy(), set(int(k) for k in kinds))
    plan.set_tuning(mode="fp32")
    return out


@pytest.mark.parametrize("dims,deg,mass", [((16, 16, 16), 5, True), ((13, 11, 9), 5, True),
                                           ((12, 16, 8), 3, True), ((16, 12, 10), 5, False)])
def test_fused_hexa_q2_equals_two_kernel_path(dims, deg, mass):
    print(f"case: fused hexa {dims} Q2 degree {deg} mass {mass}", flush=True)
    m = _hexa(*dims)
    plan, eta = _plan(m, 2, deg, mass)
    out = _fused_and_split(plan, eta)
    fused, kf = out[-1]
    split, ks = out[8]
    print(f"hexa {dims} Q2 degree {deg} mass {mass}: kinds fused {kf}, split {ks}")
    assert 3 in kf and 3 not in ks          # the fused kernel ran, then not
    assert 3 not in out[0][1]               # nor by default (autotuner off)
    assert np.array_equal(out[0][0], split)
    assert np.isfinite(fused).all()
    assert np.array_equal(fused, split)


@pytest.mark.parametrize("n,p,deg", [(48, 3, 6), (37, 4, 8)])
def test_fused_quad_high_order_equals_two_kernel_path(n, p, deg):
    print(f"case: fused quad {n}^2 Q{p} degree {deg}", flush=True)
    m = _quad(n)
    plan, eta = _plan(m, p, deg)
    out = _fused_and_split(plan, eta)
    fused, kf = out[-1]
    split, _ = out[8]
    assert 3 in kf
    assert np.isfinite(fused).all() and np.array_equal(fused, split)


def test_fused_vs_oracle_and_repeatable():
    import torch
    print("case: fused hexa 8^3 Q2 vs oracle, repeated", flush=True)
    m = _hexa(8, 8, 8, seed=4)
    plan, eta = _plan(m, 2, 5)
    ref = O.assemble_bilinear(_om(m), 2, ["stiffness", "mass"], eta=eta, degree=5)
    colptr, rowval = plan.csc_structure()
    assert np.array_equal(colptr + 1, ref[0]) and np.array_equal(rowval + 1, ref[1])
    eta_d = torch.from_numpy(eta).cuda()
    plan.set_tuning(elem_waves=-1, mode="fp32")
    vals = []
    for _ in range(3):
        v = torch.empty(plan.nnz, dtype=torch.float64, device="cuda")
        plan.assemble_device(v.data_ptr(), [eta_d.data_ptr(), None], "fp32")
        plan.V.mesh.backend.sync()
        vals.append(v.cpu().numpy())
    assert all(np.array_equal(vals[0], x) for x in vals[1:])
    err = np.abs(vals[0] - ref[2]).max() / np.abs(ref[2]).max()
    print(f"fused 8^3 vs oracle: {err:.2e}")
    assert 0 < err <= 1e-6
