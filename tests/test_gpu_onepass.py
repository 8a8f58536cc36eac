"""GPU parity of the one-pass structured Q2 kernel (bh_sf.hip): continuous Q2
stiffness + mass on structured hexahedral grids, FP64 element matrices by sum
factorisation summed per tile in LDS, every CSR row written once.

Tolerance: every matrix entry within 1e-12 * max|A_oracle| of the oracle's
assembly of the same mesh and coefficients, the CSR pattern bit-exact.  Grids
whose sides are not multiples of the 4-cell tile, single-layer z segments
(every segment then re-runs the layer below it), constant and per-cell
coefficients, stiffness or mass alone, Hexa8 and curved Hexa27 cells.  At the
full 64^3 size: equal to the element + row-gather path to 1e-13 (the two sum
each entry in different fixed orders) and bitwise repeatable.
"""
import numpy as np
import pytest

import bcubehip as B
from oracle import oracle as O

pytestmark = pytest.mark.gpu
RTOL = 1e-12


def _om(m):
    return O.OMesh(m.dim, m.coords, m.celltypes, m.c2n_ptr, m.c2n)


def _grid(nx, ny, nz, seed=0, curved=False):
    if curved:
        m = B.hexa27_mesh(nx + 1, ny + 1, nz + 1)
        B.sinusoidal_map(m, 0.05, seed=seed)
    else:
        m = B.hexa_mesh(nx + 1, ny + 1, nz + 1)
        B.jitter_nodes(m, 0.15, 1.0 / max(nx, ny, nz), seed=seed)
    return m


def _plan(m, terms, eta=None, c=1.0):
    U = B.TrialFESpace(B.FunctionSpace("Lagrange", 2), m)
    dΩ = B.Measure(B.CellDomain(m), 5)
    ηc = B.MeshCellData(eta) if eta is not None else 1.7

    def a(u, v):
        parts = []
        if "stiffness" in terms:
            parts.append(B.integrate(ηc * B.dot(B.grad(u), B.grad(v)), dΩ))
        if "mass" in terms:
            parts.append(B.integrate(c * u * v, dΩ))
        out = parts[0]
        for p in parts[1:]:
            out = out + p
        return out

    return B.BilinearPlan(a, U, B.TestFESpace(U))


def _oracle(m, terms, eta=None, cst=None):
    """The oracle's matrix of the same terms (stiffness with η per cell or the
    constant cst[0], mass scaled by cst[1])."""
    total = None
    for t in terms:
        e = eta if (t == "stiffness" and eta is not None) else None
        colptr, rowval, nz, n = O.assemble_bilinear(_om(m), 2, [t], eta=e, degree=5)
        if cst is not None:
            nz = nz * (cst[0] if t == "stiffness" and eta is None else 1.0)
            nz = nz * (cst[1] if t == "mass" else 1.0)
        total = nz if total is None else total + nz
    return colptr, rowval, total


def _timed_kinds(plan):
    backend = plan.V.mesh.backend
    backend.set_timing(True)
    try:
        vals = plan.assemble_host()
        ms, kinds = backend.timings_ex()
    finally:
        backend.set_timing(False)
    return vals, list(kinds)


@pytest.mark.parametrize("dims,terms", [((2, 3, 2), ("stiffness", "mass")),
                                        ((5, 6, 3), ("stiffness",)),
                                        ((9, 5, 4), ("stiffness", "mass")),
                                        ((4, 4, 4), ("mass",)),
                                        ((7, 9, 6), ("stiffness", "mass"))])
def test_onepass_vs_oracle(dims, terms):
    m = _grid(*dims, seed=sum(dims))
    eta = np.random.default_rng(3).uniform(0.5, 2.0, m.ncells)
    plan = _plan(m, terms, eta)
    info = plan.onepass()
    assert info["ok"] and (info["nx"], info["ny"], info["nz"]) == dims, info
    vals, kinds = _timed_kinds(plan)
    assert kinds == [3], kinds  # the one-pass kernel, nothing else
    colptr, rowval, ref = _oracle(m, terms, eta)
    c, r = plan.csc_structure()
    assert np.array_equal(c + 1, colptr) and np.array_equal(r + 1, rowval)
    err = np.abs(vals - ref).max() / np.abs(ref).max()
    print(f"{dims} {terms}: layers/group {info['layers_per_group']}, "
          f"{info['templates']} templates, rel err {err:.2e}")
    assert err <= RTOL


@pytest.mark.parametrize("dims,curved", [((5, 6, 3), False), ((9, 5, 4), False), ((4, 4, 4), True)])
def test_onepass_fp32_vs_oracle(dims, curved):
    """The FP32 variant (shape -2 in BH_FP32): FP32 element matrices and row sums
    from FP64 geometry, within 1e-6 of max|A| of the FP64 oracle."""
    m = _grid(*dims, seed=7, curved=curved)
    eta = np.random.default_rng(4).uniform(0.5, 2.0, m.ncells)
    plan = _plan(m, ("stiffness", "mass"), eta)
    plan.set_tuning(-2, mode="fp32")
    backend = m.backend
    backend.set_timing(True)
    try:
        vals = plan.assemble_host(mode="fp32")
        _, kinds = backend.timings_ex()
    finally:
        backend.set_timing(False)
    assert list(kinds) == [3]
    colptr, rowval, ref = _oracle(m, ("stiffness", "mass"), eta)
    err = np.abs(vals - ref).max() / np.abs(ref).max()
    print(f"fp32 one pass {dims} curved={curved}: rel err {err:.2e}")
    assert 0 < err <= 1e-6
    assert np.array_equal(vals, plan.assemble_host(mode="fp32"))


def test_onepass_constant_coefficients():
    m = _grid(6, 5, 7, seed=4)
    plan = _plan(m, ("stiffness", "mass"), None, c=0.3)
    vals = plan.assemble_host()
    colptr, rowval, ref = _oracle(m, ("stiffness", "mass"), None, cst=(1.7, 0.3))
    assert np.abs(vals - ref).max() <= RTOL * np.abs(ref).max()


@pytest.mark.parametrize("n", [3, 5])
def test_onepass_curved_hexa27_vs_oracle(n):
    m = _grid(n, n, n, seed=1, curved=True)
    eta = np.random.default_rng(2).uniform(0.5, 2.0, m.ncells)
    plan = _plan(m, ("stiffness", "mass"), eta)
    assert plan.onepass()["ok"]
    vals, kinds = _timed_kinds(plan)
    assert kinds == [3]
    colptr, rowval, ref = _oracle(m, ("stiffness", "mass"), eta)
    err = np.abs(vals - ref).max() / np.abs(ref).max()
    print(f"curved Hexa27 {n}^3: rel err {err:.2e}")
    assert err <= RTOL


def test_onepass_matches_gather_path_and_repeats():
    """12^3: the one-pass values against the element + row-gather kernels
    (shape 8), bitwise repeatable, and the device-coefficient call equal to the
    host one."""
    import torch

    m = _grid(12, 12, 12, seed=9)
    eta = np.random.default_rng(5).uniform(0.5, 2.0, m.ncells)
    plan = _plan(m, ("stiffness", "mass"), eta)
    one = plan.assemble_host().copy()
    assert np.array_equal(one, plan.assemble_host())
    dev = torch.device("cuda:0")
    out = torch.zeros(plan.nnz, dtype=torch.float64, device=dev)
    eta_d = torch.from_numpy(eta).to(dev)
    plan.assemble_device(out.data_ptr(), [eta_d.data_ptr(), None])
    m.backend.sync()
    assert np.array_equal(out.cpu().numpy(), one)
    plan.set_tuning(8, mode="fp64")
    two = plan.assemble_host()
    err = np.abs(one - two).max() / np.abs(two).max()
    print(f"one-pass vs element + row gather: {err:.2e}")
    assert err <= 1e-13
    assert plan.tuning("fp64")["elem_waves"] == 8


def test_onepass_not_taken_where_ineligible():
    """Q1 on hexahedra, Q2 on quads and a cell numbering that is not a grid keep
    the element + row-gather kernels, and say why."""
    m = B.hexa_mesh(4, 4, 4)
    U = B.TrialFESpace(B.FunctionSpace("Lagrange", 1), m)
    dΩ = B.Measure(B.CellDomain(m), 2)
    p = B.BilinearPlan(lambda u, v: B.integrate(u * v, dΩ), U, B.TestFESpace(U))
    info = p.onepass()
    assert not info["ok"] and "Q2" in info["why"]
    _, kinds = _timed_kinds(p)
    assert kinds == [0, 1]
    # the same grid with its cells listed in reverse: the dof map is still a
    # lattice, but cell c and c + 1 are not x-neighbours in the grid order
    g = B.hexa_mesh(4, 4, 4)
    c2n = np.asarray(g.c2n).reshape(-1, 8)[::-1].reshape(-1)
    r = B.Mesh(g.nodes(), ["Hexa8"] * g.ncells, list(c2n.reshape(-1, 8)))
    U = B.TrialFESpace(B.FunctionSpace("Lagrange", 2), r)
    dΩ = B.Measure(B.CellDomain(r), 5)
    p = B.BilinearPlan(lambda u, v: B.integrate(u * v, dΩ), U, B.TestFESpace(U))
    ok = p.onepass()["ok"]
    vals, kinds = _timed_kinds(p)
    colptr, rowval, nz, n = O.assemble_bilinear(_om(r), 2, ["mass"], degree=5)
    assert np.abs(vals - nz).max() <= RTOL * np.abs(nz).max()
    assert kinds == ([3] if ok else [0, 1])


def test_onepass_full_size_matches_gather_path():
    """64^3 Q2 (135,005,697 nnz): equal to the element + row-gather kernels to
    1e-13 of max|A|, and bitwise repeatable."""
    import torch

    n = 64
    m = _grid(n, n, n, seed=0)
    eta = np.random.default_rng(1).uniform(0.5, 2.0, m.ncells)
    plan = _plan(m, ("stiffness", "mass"), eta)
    assert plan.nnz == 135_005_697
    info = plan.onepass()
    assert info["ok"] and info["nx"] == 64, info
    dev = torch.device("cuda:0")
    eta_d = torch.from_numpy(eta).to(dev)
    a = torch.zeros(plan.nnz, dtype=torch.float64, device=dev)
    b = torch.zeros_like(a)
    plan.assemble_device(a.data_ptr(), [eta_d.data_ptr(), None])
    plan.assemble_device(b.data_ptr(), [eta_d.data_ptr(), None])
    m.backend.sync()
    assert torch.equal(a, b)
    plan.set_tuning(8, mode="fp64")
    plan.assemble_device(b.data_ptr(), [eta_d.data_ptr(), None])
    m.backend.sync()
    scale = b.abs().max().item()
    err = (a - b).abs().max().item() / scale
    print(f"64^3 one-pass vs element + row gather: {err:.2e}")
    assert err <= 1e-13
    plan.set_tuning(-2, mode="fp32")
    plan.assemble_device(b.data_ptr(), [eta_d.data_ptr(), None], "fp32")
    m.backend.sync()
    e32 = (a - b).abs().max().item() / scale
    print(f"64^3 FP32 one pass vs FP64: {e32:.2e}")
    assert 0 < e32 <= 1e-6
