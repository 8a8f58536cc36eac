"""GPU parity of the COVO Euler DG residual (benchmark/covo.jl,
driver_bench_covo.jl): the volume flux, the Roe flux on interior and periodic
faces and the explicit step, assembled by bh_euler_* through assemble_linear
on the reference's MultiFESpace(ρ, ρu, ρE, ϕ) of DG Q1 spaces, against the
oracle's restatement (FP64 within 1e-12 of max|b|), plus free stream and
conservation at larger sizes and bitwise repeatability."""
import numpy as np
import pytest

import bcubehip as B
from oracle import oracle as O

pytestmark = pytest.mark.gpu
GAMMA = 1.4
RTOL = 1e-12
L_HALF = 0.05
TAGS = ("West", "East", "South", "North")


def _meshes(n, jitter=0.0, seed=0, rotate=False):
    """The same n x n Quad4 mesh of [-l, l]² as a product Mesh (with boundary
    tags) and an oracle mesh; optional interior jitter and rotated node lists."""
    t = np.linspace(-L_HALF, L_HALF, n + 1)
    X = np.array([[x, y] for y in t for x in t])
    if jitter:
        rng = np.random.default_rng(seed)
        inner = np.all((np.abs(X + L_HALF) > 1e-12) & (np.abs(X - L_HALF) > 1e-12), axis=1)
        X[inner] += rng.uniform(-jitter, jitter, (inner.sum(), 2)) * 2 * L_HALF / n
    cells = []
    for j in range(n):
        for i in range(n):
            a = j * (n + 1) + i
            c = [a, a + 1, a + n + 2, a + n + 1]
            if rotate and (i + j) % 2:
                k = (i + 2 * j) % 3 + 1
                c = c[k:] + c[:k]
            cells.append(c)
    cells = np.array(cells, dtype=np.int64)
    m = B.Mesh(X, ["Quad4"] * (n * n), cells)
    gx, gy = np.meshgrid(np.arange(n + 1), np.arange(n + 1), indexing="xy")
    gx, gy = gx.ravel(), gy.ravel()
    m.bc_nodes = {"West": np.flatnonzero(gx == 0), "East": np.flatnonzero(gx == n),
                  "South": np.flatnonzero(gy == 0), "North": np.flatnonzero(gy == n)}
    om = O.OMesh(2, X.reshape(-1), np.zeros(n * n, dtype=np.int32),
                 np.arange(n * n + 1, dtype=np.int64) * 4, cells.reshape(-1))
    return m, om


def _setup(m):
    fs = B.FunctionSpace("Lagrange", 1)
    U_sca = B.TrialFESpace(fs, m, "discontinuous", size=1)
    U_vec = B.TrialFESpace(fs, m, "discontinuous", size=2)
    V_sca, V_vec = B.TestFESpace(U_sca), B.TestFESpace(U_vec)
    U = B.MultiFESpace(U_sca, U_vec, U_sca, U_sca)
    V = B.MultiFESpace(V_sca, V_vec, V_sca, V_sca)
    return U, V


def _forms(m, U, values, fcoef=-1.0):
    u = B.FEFunction(U, values)
    dΩ = B.Measure(B.CellDomain(m), 3)
    dΓ = B.Measure(B.InteriorFaceDomain(m), 3)
    px = B.PeriodicBCType(B.Translation([-2 * L_HALF, 0.0]), ("East",), ("West",))
    py = B.PeriodicBCType(B.Translation([0.0, 2 * L_HALF]), ("South",), ("North",))
    dΓx = B.Measure(B.BoundaryFaceDomain(m, px), 3)
    dΓy = B.Measure(B.BoundaryFaceDomain(m, py), 3)
    nΓ, nx, ny = (B.get_face_normals(d) for d in (dΓ, dΓx, dΓy))

    def flux_Ω(u, v):
        return B.compose(B.EulerFlux(GAMMA), u, tuple(map(B.grad, v)))

    def flux_Γ(u, v, n):
        return B.compose(B.RoeFlux(GAMMA), B.side_n(u), B.side_p(u), B.jump(v), B.side_n(n))

    forms = {
        "l": lambda v: (B.integrate(flux_Ω(u, v), dΩ) - B.integrate(flux_Γ(u, v, nΓ), dΓ)
                        - B.integrate(flux_Γ(u, v, nx), dΓx) - B.integrate(flux_Γ(u, v, ny), dΓy)),
        "l_vol": lambda v: B.integrate(flux_Ω(u, v), dΩ),
        "l_Γ": lambda v: B.integrate(flux_Γ(u, v, nΓ), dΓ),
    }
    return forms, (dΓx, dΓy)


def _state(nc, seed, smooth=None):
    rng = np.random.default_rng(seed)
    rho = rng.uniform(0.8, 1.2, (nc, 4))
    vel = rng.uniform(-0.4, 0.4, (2, nc, 4))
    p = rng.uniform(0.7, 1.3, (nc, 4))
    rE = p / (GAMMA - 1) + 0.5 * rho * (vel ** 2).sum(0)
    phi = rng.uniform(0.0, 1.0, (nc, 4))
    return np.stack([rho, rho * vel[0], rho * vel[1], rE, phi])


def _to_global(U, Uo):
    f = B.assembler.Fields(U)
    u = np.zeros(f.ndofs)
    u[f.maps] = Uo.reshape(5, -1)
    return u, f.maps


@pytest.mark.parametrize("n,jitter,rotate", [(4, 0.0, False), (6, 0.2, True)])
def test_euler_residual_vs_oracle(n, jitter, rotate):
    m, om = _meshes(n, jitter, seed=n, rotate=rotate)
    U, V = _setup(m)
    Uo = _state(n * n, seed=n)
    u, maps = _to_global(U, Uo)
    forms, (dΓx, dΓy) = _forms(m, U, u)
    per = np.concatenate([dΓx.domain.periodic_faces(), dΓy.domain.periodic_faces()])
    ref_per = (O.periodic_faces(om, m.bc_nodes, np.array([-2 * L_HALF, 0.0]), ("East",), ("West",))
               + O.periodic_faces(om, m.bc_nodes, np.array([0.0, 2 * L_HALF]), ("South",),
                                  ("North",)))
    assert np.array_equal(per, np.array(ref_per, dtype=np.int32))
    print(f"Euler n={n} jitter={jitter} rotate={rotate}", flush=True)
    for name, terms, fcoef in (("l", ("vol", "face"), -1.0), ("l_vol", ("vol",), 0.0),
                               ("l_Γ", ("face",), 1.0)):
        b = B.assemble_linear(forms[name], V)
        if name == "l_Γ":
            ref = -O.euler_dg_residual(om, Uo, GAMMA, 3, (), terms=("face",))
        else:
            ref = O.euler_dg_residual(om, Uo, GAMMA, 3, ref_per if name == "l" else (),
                                      terms=terms)
        got = b[maps].reshape(5, n * n, 4)
        err = np.abs(got - ref).max() / np.abs(ref).max()
        print(f"  {name}: rel err {err:.2e}")
        assert err <= RTOL


def test_euler_explicit_step_vs_oracle():
    n = 5
    m, om = _meshes(n, 0.15, seed=9, rotate=True)
    U, V = _setup(m)
    Uo = _state(n * n, seed=1)
    u, maps = _to_global(U, Uo)
    _, (dΓx, dΓy) = _forms(m, U, u)
    plan = B.EulerDG(V, GAMMA, 3, (dΓx.domain, dΓy.domain))
    dt = 1e-4
    out = plan.step(u, dt)
    per = [list(r) for r in np.concatenate([dΓx.domain.periodic_faces(),
                                            dΓy.domain.periodic_faces()])]
    R = O.euler_dg_residual(om, Uo, GAMMA, 3, per)
    ref = Uo + dt * O.euler_dg_mass_inverse_apply(om, R, 3)
    got = out[maps].reshape(5, n * n, 4)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    inc = np.abs((got - Uo) - (ref - Uo)).max() / np.abs(ref - Uo).max()
    print(f"step: rel err {err:.2e}, increment rel err {inc:.2e}")
    assert inc <= 1e-11


def test_euler_free_stream_and_conservation_at_64():
    n = 64
    m, _ = _meshes(n, 0.2, seed=3, rotate=True)
    U, V = _setup(m)
    w = np.array([1.1, 0.3, -0.2, 2.9, 0.4])
    u, maps = _to_global(U, np.broadcast_to(w[:, None, None], (5, n * n, 4)))
    forms, _ = _forms(m, U, u)
    b = B.assemble_linear(forms["l"], V)
    vol = B.assemble_linear(forms["l_vol"], V)
    print(f"free stream 64²: max|b| {np.abs(b).max():.2e} vs volume {np.abs(vol).max():.2e}")
    assert np.abs(b).max() <= 1e-12 * np.abs(vol).max()
    u2, _ = _to_global(U, _state(n * n, seed=5))
    forms2, _ = _forms(m, U, u2)
    b2 = B.assemble_linear(forms2["l"], V)
    sums = b2[maps].sum(axis=1)
    assert np.abs(sums).max() <= 1e-11 * np.abs(b2).max()


def test_euler_device_path_bitwise_repeatable():
    import torch
    n = 16
    m, _ = _meshes(n, 0.1, seed=4)
    U, V = _setup(m)
    u, _ = _to_global(U, _state(n * n, seed=2))
    _, (dΓx, dΓy) = _forms(m, U, u)
    plan = B.EulerDG(V, GAMMA, 3, (dΓx.domain, dΓy.domain))
    host = plan.residual(u)
    du = torch.from_numpy(u).cuda()
    db = torch.zeros_like(du)
    for _ in range(2):
        db.zero_()
        torch.cuda.synchronize()
        plan.residual_device(du.data_ptr(), db.data_ptr())
        B.lib().bh_sync(m.backend.ctx)
        assert np.array_equal(db.cpu().numpy(), host)
    step_host = plan.step(u, 1e-5)
    do = torch.empty_like(du)
    plan.step_device(du.data_ptr(), do.data_ptr(), 1e-5)
    B.lib().bh_sync(m.backend.ctx)
    assert np.array_equal(do.cpu().numpy(), step_host)


def test_euler_steps_in_a_hip_graph_replay_bitwise():
    """covo.jl's time loop as a captured graph: two explicit steps (u -> v ->
    u) captured once and replayed; every replay equals the eager steps bit for
    bit (eager work and synchronisations between replays)."""
    import torch
    n = 32
    m, _ = _meshes(n, 0.1, seed=6)
    U, V = _setup(m)
    u0, _ = _to_global(U, _state(n * n, seed=8))
    _, (dΓx, dΓy) = _forms(m, U, u0)
    plan = B.EulerDG(V, GAMMA, 3, (dΓx.domain, dΓy.domain))
    dt = 2e-5
    # eager reference trajectory on the host path
    ref = [u0]
    for _ in range(6):
        ref.append(plan.step(ref[-1], dt))
    stream = torch.cuda.Stream()
    m.backend.set_stream(stream.cuda_stream)
    a = torch.from_numpy(u0).cuda()
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    with m.backend.capture() as g:
        plan.step_device(a.data_ptr(), b.data_ptr(), dt)
        plan.step_device(b.data_ptr(), a.data_ptr(), dt)
    for r in range(3):
        g.launch()
        B.lib().bh_sync(m.backend.ctx)
        got = a.cpu().numpy()
        assert np.array_equal(got, ref[2 * (r + 1)]), r
        _ = (a * 2.0).sum().item()   # eager work between replays
    m.backend.set_stream(None)


def test_euler_residual_soa_numbering():
    """MultiFESpace(...; arrayOfStruct=false): a cell's dofs spread over the
    whole vector (space by space); the kernels read them through the maps."""
    n = 5
    m, om = _meshes(n, 0.1, seed=11, rotate=True)
    fs = B.FunctionSpace("Lagrange", 1)
    U_sca = B.TrialFESpace(fs, m, "discontinuous", size=1)
    U_vec = B.TrialFESpace(fs, m, "discontinuous", size=2)
    U = B.MultiFESpace(U_sca, U_vec, U_sca, U_sca, arrayOfStruct=False)
    V = B.MultiFESpace(B.TestFESpace(U_sca), B.TestFESpace(U_vec), B.TestFESpace(U_sca),
                       B.TestFESpace(U_sca), arrayOfStruct=False)
    Uo = _state(n * n, seed=12)
    u, maps = _to_global(U, Uo)
    assert not all(np.array_equal(np.sort(maps[:, 4 * c:4 * c + 4].ravel()),
                                  np.arange(20 * c, 20 * c + 20)) for c in range(n * n))
    forms, (dΓx, dΓy) = _forms(m, U, u)
    b = B.assemble_linear(forms["l"], V)
    per = [list(r) for r in np.concatenate([dΓx.domain.periodic_faces(),
                                            dΓy.domain.periodic_faces()])]
    ref = O.euler_dg_residual(om, Uo, GAMMA, 3, per)
    err = np.abs(b[maps].reshape(5, n * n, 4) - ref).max() / np.abs(ref).max()
    print(f"SoA numbering: rel err {err:.2e}")
    assert err <= RTOL


def _mixed_meshes(n, jitter, seed):
    """n x n cells of [-l, l]², every third cell split into two triangles."""
    t = np.linspace(-L_HALF, L_HALF, n + 1)
    X = np.array([[x, y] for y in t for x in t])
    rng = np.random.default_rng(seed)
    inner = np.all((np.abs(X + L_HALF) > 1e-12) & (np.abs(X - L_HALF) > 1e-12), axis=1)
    X[inner] += rng.uniform(-jitter, jitter, (inner.sum(), 2)) * 2 * L_HALF / n
    cells = []
    for j in range(n):
        for i in range(n):
            a = j * (n + 1) + i
            if (i + 2 * j) % 3 == 0:
                cells += [[a, a + 1, a + n + 2], [a, a + n + 2, a + n + 1]]
            else:
                cells.append([a, a + 1, a + n + 2, a + n + 1])
    types = ["Tri3" if len(c) == 3 else "Quad4" for c in cells]
    m = B.Mesh(X, types, cells)
    gx, gy = np.meshgrid(np.arange(n + 1), np.arange(n + 1), indexing="xy")
    gx, gy = gx.ravel(), gy.ravel()
    m.bc_nodes = {"West": np.flatnonzero(gx == 0), "East": np.flatnonzero(gx == n),
                  "South": np.flatnonzero(gy == 0), "North": np.flatnonzero(gy == n)}
    om = O.OMesh(2, X.reshape(-1), m.celltypes, m.c2n_ptr, m.c2n)
    return m, om


def _to_global_mixed(U, Uo, m):
    """u[maps[k][dof(c, l)]] = Uo[k, c, l] for the l < 3 or 4 dofs of each cell."""
    from bcubehip.assembler import Fields
    f = Fields(U)
    dofs = _dofs(f.scalar)
    nl = dofs.size // m.ncells
    u = np.zeros(f.ndofs)
    idx = np.full((5, m.ncells, 4), -1, dtype=np.int64)
    for c in range(m.ncells):
        for l in range(3 if m.celltypes[c] == 2 else 4):
            g = dofs[c * nl + l]
            idx[:, c, l] = f.maps[:, g]
            u[f.maps[:, g]] = Uo[:, c, l]
    return u, idx


def _dofs(space):
    import ctypes as C
    from bcubehip import _lib as L
    nl, nd = C.c_int32(), C.c_int64()
    L.check(L.lib().bh_space_info(space.handle, C.byref(nd), C.byref(nl)))
    out = np.zeros(space.mesh.ncells * nl.value, dtype=np.int32)
    L.check(L.lib().bh_space_dofs(space.handle, out.ctypes.data))
    return out


def test_euler_on_a_mixed_tri_quad_mesh():
    """The residual and the step on Quad4 + Tri3 cells (COVO's MeshConfig
    triquad): P1 triangles with the degree-3 triangle rule, against the
    oracle; free stream on the same mesh."""
    n = 6
    m, om = _mixed_meshes(n, 0.15, seed=13)
    assert (m.celltypes == 2).any() and (m.celltypes == 0).any()
    U, V = _setup(m)
    nc = m.ncells
    Uo = _state(nc, seed=14)
    Uo[:, m.celltypes == 2, 3] = 0.0
    u, idx = _to_global_mixed(U, Uo, m)
    forms, (dΓx, dΓy) = _forms(m, U, u)
    b = B.assemble_linear(forms["l"], V)
    per = [list(r) for r in np.concatenate([dΓx.domain.periodic_faces(),
                                            dΓy.domain.periodic_faces()])]
    ref_per = (O.periodic_faces(om, m.bc_nodes, np.array([-2 * L_HALF, 0.0]), ("East",), ("West",))
               + O.periodic_faces(om, m.bc_nodes, np.array([0.0, 2 * L_HALF]), ("South",),
                                  ("North",)))
    assert per == [list(r) for r in ref_per]
    ref = O.euler_dg_residual(om, Uo, GAMMA, 3, per)
    mask = idx >= 0
    got = np.zeros_like(ref)
    got[mask] = b[idx[mask]]
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print(f"mixed mesh residual: rel err {err:.2e}", flush=True)
    assert err <= RTOL
    plan = B.EulerDG(V, GAMMA, 3, (dΓx.domain, dΓy.domain))
    dt = 1e-4
    out = plan.step(u, dt)
    refs = Uo + dt * O.euler_dg_mass_inverse_apply(om, ref, 3)
    gs = np.zeros_like(refs)
    gs[mask] = out[idx[mask]]
    inc = np.abs((gs - Uo)[mask] - (refs - Uo)[mask]).max() / np.abs((refs - Uo)[mask]).max()
    print(f"mixed mesh step: increment rel err {inc:.2e}")
    assert inc <= 1e-11
    w = np.array([1.1, 0.3, -0.2, 2.9, 0.4])
    uf, _ = _to_global_mixed(U, np.broadcast_to(w[:, None, None], (5, nc, 4)).copy(), m)
    ff, _ = _forms(m, U, uf)
    bf = B.assemble_linear(ff["l"], V)
    vol = B.assemble_linear(ff["l_vol"], V)
    assert np.abs(bf).max() <= 1e-12 * np.abs(vol).max()


def test_euler_and_face_term_error_paths():
    """Refusals, not faults: a continuous space, a 3D mesh for face terms,
    b aliasing u, and a periodic face on a side that already has a face."""
    import ctypes as C
    from bcubehip import _lib as L
    m, _ = _meshes(4)
    Uc = B.TrialFESpace(B.FunctionSpace("Lagrange", 1), m)
    maps = np.arange(5 * Uc.get_ndofs(), dtype=np.int64)
    h = C.c_void_p()
    rc = L.lib().bh_euler_create(Uc.handle, 3, 1.4, 0, None, maps.ctypes.data, maps.size,
                                 C.byref(h))
    assert rc == -2 and not h.value          # BH_ERR_UNSUPPORTED
    Ud = B.TrialFESpace(B.FunctionSpace("Lagrange", 1), m, "discontinuous")
    ok = np.array([[0, 0, 12, 2, 1]], dtype=np.int32)   # bottom of cell 0, top of cell 12
    per = np.array([[0, 1, 5, 3, 0]], dtype=np.int32)   # side 1 of cell 0 is interior
    maps = np.arange(5 * Ud.get_ndofs(), dtype=np.int64)
    rc = L.lib().bh_euler_create(Ud.handle, 3, 1.4, 1, per.ctypes.data, maps.ctypes.data,
                                 maps.size, C.byref(h))
    assert rc == -1 and not h.value          # BH_ERR_ARG
    rc = L.lib().bh_euler_create(Ud.handle, 3, 1.4, 1, ok.ctypes.data, maps.ctypes.data,
                                 maps.size + 7, C.byref(h))
    assert rc == -1 and not h.value          # maps not onto ndofs_total
    rc = L.lib().bh_euler_create(Ud.handle, 3, 1.4, 1, ok.ctypes.data, maps.ctypes.data,
                                 maps.size, C.byref(h))
    assert rc == 0
    u = np.ones(maps.size)
    assert L.lib().bh_euler_residual(h, 3, -1.0, u.ctypes.data, u.ctypes.data, L.HOST) == -1
    L.lib().bh_euler_destroy(h)
    Ue, Ve = _setup(m)
    u0, _ = _to_global(Ue, _state(16, seed=1))
    _, (dΓx, dΓy) = _forms(m, Ue, u0)
    plan = B.EulerDG(Ve, GAMMA, 3, (dΓx.domain, dΓy.domain))
    with pytest.raises(ValueError):          # a short state: refused before any copy
        plan.step(u0[:-1], 1e-5)
    mh = B.hexa_mesh(3, 3, 3)
    Uh = B.TrialFESpace(B.FunctionSpace("Lagrange", 1), mh, "discontinuous")
    dΓ = B.Measure(B.InteriorFaceDomain(mh), 3)
    with pytest.raises(B.BcubeHipError):
        B.assemble_bilinear(lambda u, v: B.integrate(B.side_n(u) * B.side_n(v), dΓ), Uh,
                            B.TestFESpace(Uh))
