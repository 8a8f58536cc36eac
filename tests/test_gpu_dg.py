"""GPU parity of the DG transport residual (volume + upwind interior faces)
against the oracle, and of the explicit RK4 loop built on it.

Tolerances: FP64 every residual entry within 1e-12 * max|r_oracle|; FP32
(FP64 geometry, FP32 fields and arithmetic) within 2e-6 * max|r_oracle|.
At the full 32^3 size the checks are size-independent: the residual sums to
zero per field (test functions sum to one and jump(1) = 0), and a constant
field has a vanishing residual in every cell away from the boundary.
"""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

import bcubehip as B
from oracle import oracle as O

pytestmark = pytest.mark.gpu
C = np.array([1.0, 0.5, 0.25])
RTOL = 1e-12
RTOL_FP32 = 2e-6


def _om(m):
    return O.OMesh(m.dim, m.coords, m.celltypes, m.c2n_ptr, m.c2n)


def _curved(n, seed=0):
    m = B.hexa27_mesh(n + 1, n + 1, n + 1)
    B.sinusoidal_map(m, 0.05, seed=seed)
    return m


def _jittered8(n, seed=4):
    m = B.hexa_mesh(n + 1, n + 1, n + 1)
    B.jitter_nodes(m, 0.15, 1.0 / n, seed=seed)
    return m


def _space(m, p):
    return B.TrialFESpace(B.FunctionSpace("Lagrange", p), m, "discontinuous")


def _fields(ndofs, nf, seed=2):
    return np.random.default_rng(seed).standard_normal((nf, ndofs))


@pytest.mark.parametrize("kind,n,p,nf", [("hexa27", 4, 1, 2), ("hexa27", 4, 2, 3),
                                         ("hexa27", 5, 3, 8), ("hexa8", 4, 3, 4),
                                         ("hexa8", 3, 2, 1)])
def test_residual_fp64_vs_oracle(kind, n, p, nf):
    m = _curved(n) if kind == "hexa27" else _jittered8(n)
    V = _space(m, p)
    plan = B.DGTransport(V, C, 2 * p + 1, "fp64")
    u = _fields(V.get_ndofs(), nf)
    r = plan.residual(u)
    ref = O.dg_transport(_om(m), p, C, u.T, 2 * p + 1).T
    scale = np.abs(ref).max()
    err = np.abs(r - ref).max()
    print(f"DG {kind} n={n} Q{p} nf={nf}: rel err {err / scale:.2e}, faces {plan.ninterior_faces}")
    assert err <= RTOL * scale


@pytest.mark.parametrize("p", [2, 3])
def test_residual_fp32_vs_oracle(p):
    m = _curved(5, seed=1)
    V = _space(m, p)
    plan = B.DGTransport(V, C, 2 * p + 1, "fp32")
    u = _fields(V.get_ndofs(), 8).astype(np.float32)
    r = plan.residual(u)
    assert r.dtype == np.float32
    ref = O.dg_transport(_om(m), p, C, u.astype(np.float64).T, 2 * p + 1).T
    err = np.abs(r - ref).max() / np.abs(ref).max()
    print(f"DG FP32 Q{p}: rel err {err:.2e}")
    assert err <= RTOL_FP32


def test_residual_device_deterministic_and_form_dispatch():
    import torch
    m = _curved(4, seed=3)
    V = _space(m, 3)
    u = _fields(V.get_ndofs(), 8)
    dΩ = B.Measure(B.CellDomain(m), 7)
    dΓ = B.Measure(B.InteriorFaceDomain(m), 7)
    nΓ = B.get_face_normals(dΓ)
    uh = B.FEFunction(V, u)
    flux = B.UpwindFlux(C)
    l = lambda v: (B.integrate(B.dot(C * uh, B.grad(v)), dΩ)
                   - B.integrate(B.compose(flux, B.side_n(uh), B.side_p(uh), B.jump(v),
                                           B.side_n(nΓ)), dΓ))
    r_form = B.assemble_linear(l, B.TestFESpace(V))
    plan = B.DGTransport(V, C, 7)
    ud = torch.from_numpy(u).cuda()
    r1 = plan.residual(ud)
    r2 = plan.residual(ud)
    torch.cuda.synchronize()
    assert torch.equal(r1, r2)
    assert np.array_equal(r1.cpu().numpy(), r_form)


def _dg_mass(m, p):
    colptr, rowval, nzval, n = O.assemble_bilinear(_om(m), p, ["mass"], degree=2 * p + 1,
                                                   continuous=False)
    return O.csc_to_scipy(colptr, rowval, nzval, n)


@pytest.mark.parametrize("kind,p", [("hexa27", 3), ("hexa8", 2)])
def test_inverse_mass_axpy_vs_sparse_solve(kind, p):
    import torch
    m = _curved(3, seed=5) if kind == "hexa27" else _jittered8(3)
    V = _space(m, p)
    plan = B.DGTransport(V, C, 2 * p + 1)
    nd = V.get_ndofs()
    r = _fields(nd, 2, seed=7)
    x1 = _fields(nd, 2, seed=8)
    x2 = _fields(nd, 2, seed=9)
    rd, x1d, x2d = (torch.from_numpy(a).cuda() for a in (r, x1, x2))
    o1, o2 = torch.empty_like(rd), torch.empty_like(rd)
    plan.minv_axpy(2, rd.data_ptr(), x1d.data_ptr(), 0.3, o1.data_ptr(), x2d.data_ptr(), -1.7,
                   o2.data_ptr())
    torch.cuda.synchronize()
    M = _dg_mass(m, p).tocsc()
    k = np.stack([spla.spsolve(M, r[f]) for f in range(2)])
    e1 = np.abs(o1.cpu().numpy() - (x1 + 0.3 * k)).max() / np.abs(0.3 * k).max()
    e2 = np.abs(o2.cpu().numpy() - (x2 - 1.7 * k)).max() / np.abs(1.7 * k).max()
    print(f"M^-1 {kind} Q{p}: rel err {e1:.2e} {e2:.2e}")
    assert e1 < 1e-11 and e2 < 1e-11


def test_rk4_matches_cpu_rk4_and_conserves_mass():
    """Three RK4 steps on the GPU (eager and replayed from a HIP graph) against
    the same steps with the oracle residual and a sparse mass solve."""
    import torch
    m = _curved(3, seed=6)
    p = 2
    V = _space(m, p)
    plan = B.DGTransport(V, C, 2 * p + 1)
    nd = V.get_ndofs()
    x = B.dg_dof_coordinates(V)
    u0 = np.stack([np.exp(-30 * ((x - 0.4) ** 2).sum(1)), np.sin(2 * np.pi * x[:, 0])])
    M = _dg_mass(m, p).tocsc()
    lu = spla.splu(M)
    om = _om(m)

    def L(u):
        r = O.dg_transport(om, p, C, u.T, 2 * p + 1).T
        return np.stack([lu.solve(r[f]) for f in range(r.shape[0])])

    dt = 0.01
    ref = u0.copy()
    for _ in range(3):
        k1 = L(ref)
        k2 = L(ref + dt / 2 * k1)
        k3 = L(ref + dt / 2 * k2)
        k4 = L(ref + dt * k3)
        ref = ref + dt / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
    ud = torch.from_numpy(u0).cuda()
    work = None
    for _ in range(3):
        work = plan.rk4_step(ud, dt, work)
    torch.cuda.synchronize()
    got = ud.cpu().numpy()
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print(f"RK4 3 steps: rel err {err:.2e}")
    assert err < 1e-11
    mass0 = np.ones(nd) @ (M @ u0.T)
    mass3 = np.ones(nd) @ (M @ got.T)
    assert np.abs(mass3 - mass0).max() < 1e-12 * np.abs(u0).sum()
    # the same three steps replayed from a graph, from the same start
    backend = m.backend
    stream = torch.cuda.Stream()
    backend.set_stream(stream.cuda_stream)
    ug = torch.from_numpy(u0).cuda()
    torch.cuda.synchronize()
    plan.rk4_step(ug, dt, work)  # warm (first use outside capture)
    backend.sync()
    with torch.cuda.stream(stream):
        ug.copy_(torch.from_numpy(u0).cuda())
    backend.sync()
    with backend.capture() as g:
        plan.rk4_step(ug, dt, work)
    for rep in range(3):
        g.launch()
        backend.sync()
    backend.set_stream(None)
    assert torch.equal(ug, ud)


@pytest.mark.parametrize("p,mode", [(2, "fp64"), (3, "fp32"), (3, "fp64"), (1, "fp32")])
def test_fused_rk4_stage_bitwise_equals_two_pass(p, mode):
    """bh_dg_transport_stage (the residual kernel applying M⁻¹ and both updates,
    or residual + update through the work array for FP64 Q3) gives the RK4
    step of four residuals and four M⁻¹ updates bit for bit; an output that
    aliases the stage input is refused."""
    import torch
    m = _curved(3, seed=8)
    V = _space(m, p)
    plan = B.DGTransport(V, C, 2 * p + 1, mode)
    dt = torch.float32 if mode == "fp32" else torch.float64
    u0 = torch.from_numpy(_fields(V.get_ndofs(), 3, seed=4)).to("cuda", dt)
    a, b = u0.clone(), u0.clone()
    wa, wb = None, None
    for _ in range(2):
        wa = plan.rk4_step(a, 0.01, wa)
        wb = plan.rk4_step_two_pass(b, 0.01, wb)
    torch.cuda.synchronize()
    assert torch.equal(a, b) and not torch.equal(a, u0)
    with pytest.raises(B.BcubeHipError):
        plan.stage(3, a.data_ptr(), wa[1].data_ptr(), 0.1, a.data_ptr())


def test_full_size_invariants_fp32():
    """The C4 workload: 32^3 curved Hexa27 cells, DG Q3, 8 fields, FP32."""
    import torch
    n = 32
    m = _curved(n)
    V = _space(m, 3)
    plan = B.DGTransport(V, C, 7, "fp32")
    assert plan.ndofs == 2_097_152 and plan.ninterior_faces == 3 * n * n * (n - 1)
    u = torch.from_numpy(_fields(plan.ndofs, 8).astype(np.float32)).cuda()
    r = plan.residual(u).double()
    s = r.sum(dim=1).abs().max().item()
    assert s < 1e-5 * r.abs().sum(dim=1).min().item()
    one = torch.ones((1, plan.ndofs), dtype=torch.float32, device="cuda")
    r1 = plan.residual(one).reshape(n, n, n, 64).cpu().numpy()[1:-1, 1:-1, 1:-1]
    r1_scale = plan.residual(one).abs().max().item()
    print(f"C4 32^3: Σr/Σ|r| {s:.2e}; constant field interior max {np.abs(r1).max():.2e} "
          f"(boundary scale {r1_scale:.2e})")
    assert np.abs(r1).max() < 1e-5 * r1_scale
