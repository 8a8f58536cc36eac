"""bh_assemble_bilinear_noalloc (the allocation-free variant of
bh_assemble_bilinear, element rows in the library's static scratch) against
bh_assemble_bilinear through the C ABI: two calls in a row, the second on a
smaller mesh, in FP64 and FP32; both entry points run the same kernels in the
same order, so the values are bitwise equal.  Also its refusals: a mesh too
large for the scratch, tables not yet on the device, a NULL output."""
import ctypes as C

import numpy as np
import pytest

import bcubehip as B
from bcubehip import _lib as L

pytestmark = pytest.mark.gpu

SCRATCH = 64 << 20


def _plan(n, backend, seed):
    m = B.hexa_mesh(n + 1, n + 1, n + 1, backend=backend)
    B.jitter_nodes(m, 0.15, 1.0 / n, seed=seed)
    eta = np.random.default_rng(seed + 1).uniform(0.5, 2.0, m.ncells)
    U = B.TrialFESpace(B.FunctionSpace("Lagrange", 2), m)
    V = B.TestFESpace(U)
    dΩ = B.Measure(B.CellDomain(m), 5)
    ηc = B.MeshCellData(eta)
    plan = B.BilinearPlan(lambda u, v: B.integrate(ηc * B.dot(B.grad(u), B.grad(v)), dΩ)
                          + B.integrate(u * v, dΩ), U, V)
    return plan, eta


def _terms(plan, eta_d):
    ts = [L.BhTerm(t.form, t.coef_kind, eta_d.data_ptr() if t.coef_kind == L.COEF_CELL else None,
                   t.scale) for t in plan.terms]
    return (L.BhTerm * len(ts))(*ts)


def test_noalloc_equals_assemble_bilinear():
    import torch
    dev = torch.device("cuda:0")
    be = B.BcubeBackendHIP(0)
    lib = L.lib()
    for mode in ("fp64", "fp32"):
        for n, seed in ((8, 3), (5, 7)):  # second call on a smaller mesh
            plan, eta = _plan(n, be, seed)
            eta_d = torch.from_numpy(eta).to(dev)
            terms = _terms(plan, eta_d)
            ref = torch.empty(plan.nnz, dtype=torch.float64, device=dev)
            rc = lib.bh_assemble_bilinear(plan.handle, plan.nterms, terms, plan.measure.degree,
                                          L.MODE[mode], ref.data_ptr(), L.DEVICE)
            L.check(rc, be.ctx, "bh_assemble_bilinear")
            got = torch.full((plan.nnz,), float("nan"), dtype=torch.float64, device=dev)
            print(f"noalloc {mode} hexa Q2 {n}^3 nnz={plan.nnz}", flush=True)
            rc = lib.bh_assemble_bilinear_noalloc(plan.handle, plan.nterms, terms,
                                                  plan.measure.degree, L.MODE[mode],
                                                  got.data_ptr())
            L.check(rc, be.ctx, "bh_assemble_bilinear_noalloc")
            be.sync()
            assert torch.equal(got, ref), (mode, n)
            assert torch.isfinite(got).all()


def test_noalloc_refusals():
    import torch
    dev = torch.device("cuda:0")
    lib = L.lib()
    be = B.BcubeBackendHIP(0)  # a fresh context: no tables on its device yet
    plan, eta = _plan(3, be, 1)
    eta_d = torch.from_numpy(eta).to(dev)
    terms = _terms(plan, eta_d)
    out = torch.empty(plan.nnz, dtype=torch.float64, device=dev)
    rc = lib.bh_assemble_bilinear_noalloc(plan.handle, plan.nterms, terms, plan.measure.degree,
                                          L.MODE["fp64"], out.data_ptr())
    assert rc == -1 and b"bh_assemble_bilinear once" in lib.bh_last_error(be.ctx)
    rc = lib.bh_assemble_bilinear_noalloc(plan.handle, plan.nterms, terms, plan.measure.degree,
                                          L.MODE["fp64"], None)
    assert rc == -1  # BH_ERR_ARG
    # 22^3 Q2 hexa in FP64: (10648 + 16) * 27 * 32 * 8 bytes of element rows > 64 MiB
    big, eta_b = _plan(22, be, 2)
    need = (big.V.mesh.ncells + 16) * 27 * 32 * 8
    assert need > SCRATCH
    eta_bd = torch.from_numpy(eta_b).to(dev)
    tb = _terms(big, eta_bd)
    ob = torch.empty(big.nnz, dtype=torch.float64, device=dev)
    rc = lib.bh_assemble_bilinear_noalloc(big.handle, big.nterms, tb, big.measure.degree,
                                          L.MODE["fp64"], ob.data_ptr())
    assert rc == -5 and b"static scratch" in lib.bh_last_error(be.ctx)
