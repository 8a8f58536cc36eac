"""The oracle's restatement of the COVO Euler DG residual (benchmark/covo.jl),
pinned by what the mathematics fixes:
  - the Roe flux equals the textbook Roe flux ½(F(uL) + F(uR))⋅n − ½|Ã|(uR − uL),
    with Ã the flux Jacobian at the Roe-averaged state built here independently
    (complex-step derivative of F⋅n, |Ã| from its eigen-decomposition);
    consistency F(u, u, n) = F(u)⋅n; the passive scalar's upwinding;
  - free stream: a constant state leaves a zero residual on a periodic mesh
    (divergence theorem; fails on a wrong normal, weight, sign, a face left
    out or a wrong pairing);
  - conservation: on a fully periodic mesh every component's residual sums
    to zero over all dofs (fails on a side's sign);
  - the DG mass matrix: M⁻¹ 1 = (4/area) 1 on square cells.
"""
import numpy as np
import pytest

from oracle import oracle as O

GAMMA = 1.4


def _flux_n(q, n, gamma=GAMMA):
    """F(q)⋅n for q = (ρ, ρu, ρv, ρE), complex-safe."""
    rho, ru, rv, rE = q
    u, v = ru / rho, rv / rho
    p = (gamma - 1) * (rE - 0.5 * (ru * u + rv * v))
    un = u * n[0] + v * n[1]
    return np.array([rho * un, ru * un + p * n[0], rv * un + p * n[1], (rE + p) * un])


def _textbook_roe(uL, uR, n, gamma=GAMMA):
    qL, qR = np.asarray(uL[:4]), np.asarray(uR[:4])
    rL, rR = qL[0], qR[0]
    velL, velR = qL[1:3] / rL, qR[1:3] / rR
    pL = (gamma - 1) * (qL[3] - 0.5 * rL * velL @ velL)
    pR = (gamma - 1) * (qR[3] - 0.5 * rR * velR @ velR)
    HL, HR = (qL[3] + pL) / rL, (qR[3] + pR) / rR
    sL, sR = np.sqrt(rL), np.sqrt(rR)
    vel = (sL * velL + sR * velR) / (sL + sR)
    H = (sL * HL + sR * HR) / (sL + sR)
    # a state with the Roe averages (density 1): ρE = ρH - p, p = (γ-1)/γ ρ (H - |u|²/2)
    p = (gamma - 1) / gamma * (H - 0.5 * vel @ vel)
    qt = np.array([1.0, vel[0], vel[1], H - p])
    h = 1e-30
    A = np.zeros((4, 4))
    for k in range(4):
        dq = qt.astype(complex)
        dq[k] += 1j * h
        A[:, k] = np.imag(_flux_n(dq, n)) / h
    lam, R = np.linalg.eig(A)
    absA = np.real(R @ np.diag(np.abs(lam)) @ np.linalg.inv(R))
    return 0.5 * (_flux_n(qL, n) + _flux_n(qR, n)) - 0.5 * absA @ (qR - qL)


def _state(rng, n=1):
    rho = rng.uniform(0.7, 1.3, n)
    vel = rng.uniform(-0.5, 0.5, (n, 2))
    p = rng.uniform(0.6, 1.4, n)
    rE = p / (GAMMA - 1) + 0.5 * rho * (vel ** 2).sum(1)
    phi = rng.uniform(0.0, 1.0, n)
    return np.stack([rho, rho * vel[:, 0], rho * vel[:, 1], rE, phi], axis=0)


def test_roe_flux_is_the_textbook_roe_flux():
    rng = np.random.default_rng(0)
    for _ in range(20):
        uL, uR = _state(rng)[:, 0], _state(rng)[:, 0]
        t = rng.uniform(0, 2 * np.pi)
        n = np.array([np.cos(t), np.sin(t)])
        F = O.roe_flux(uL, uR, n, GAMMA)
        ref = _textbook_roe(uL, uR, n)
        assert np.abs(F[:4] - ref).max() <= 1e-11 * np.abs(ref).max()
        # consistency and the passive scalar's upwinding
        Fc = O.roe_flux(uL, uL, n, GAMMA)
        assert np.abs(Fc[:4] - _flux_n(uL[:4], n)).max() <= 1e-13 * np.abs(Fc).max()
        v12 = 0.5 * ((uL[1] / uL[0] + uR[1] / uR[0]) * n[0] + (uL[2] / uL[0] + uR[2] / uR[0]) * n[1])
        assert F[4] == pytest.approx(v12 * (uL[4] if v12 > 0 else uR[4]), rel=1e-14)


def _periodic_mesh(n, jitter=0.0, seed=0, rotate=False, l=0.05, split=False):
    t = np.linspace(-l, l, n + 1)
    X = np.array([[x, y] for y in t for x in t])
    if jitter:
        rng = np.random.default_rng(seed)
        inner = np.all((np.abs(X + l) > 1e-12) & (np.abs(X - l) > 1e-12), axis=1)
        X[inner] += rng.uniform(-jitter, jitter, (inner.sum(), 2)) * 2 * l / n
    cells = []
    for j in range(n):
        for i in range(n):
            a = j * (n + 1) + i
            c = [a, a + 1, a + n + 2, a + n + 1]
            if split and (i + 2 * j) % 3 == 0:   # two triangles, counter-clockwise
                cells += [[a, a + 1, a + n + 2], [a, a + n + 2, a + n + 1]]
                continue
            if rotate and (i + j) % 2:
                k = (i + 2 * j) % 3 + 1
                c = c[k:] + c[:k]
            cells.append(c)
    ct = np.array([O.CT["Tri3"] if len(c) == 3 else O.CT["Quad4"] for c in cells], dtype=np.int32)
    ptr = np.concatenate([[0], np.cumsum([len(c) for c in cells])]).astype(np.int64)
    m = O.OMesh(2, X.reshape(-1), ct, ptr, np.concatenate([np.array(c) for c in cells]).astype(np.int64))
    gx, gy = np.meshgrid(np.arange(n + 1), np.arange(n + 1), indexing="xy")
    gx, gy = gx.ravel(), gy.ravel()
    tags = {"West": np.flatnonzero(gx == 0), "East": np.flatnonzero(gx == n),
            "South": np.flatnonzero(gy == 0), "North": np.flatnonzero(gy == n)}
    per = (O.periodic_faces(m, tags, np.array([-2 * l, 0.0]), ("East",), ("West",))
           + O.periodic_faces(m, tags, np.array([0.0, 2 * l]), ("South",), ("North",)))
    return m, per


@pytest.mark.parametrize("jitter,rotate,split", [(0.0, False, False), (0.2, True, False),
                                                 (0.15, True, True)])
def test_free_stream_leaves_zero_residual(jitter, rotate, split):
    n = 5
    m, per = _periodic_mesh(n, jitter, seed=1, rotate=rotate, split=split)
    nc = len(m.celltypes)
    assert len(per) == 2 * n
    w = _state(np.random.default_rng(3))[:, 0]
    U = np.broadcast_to(w[:, None, None], (5, nc, 4)).copy()
    vol = O.euler_dg_residual(m, U, GAMMA, 3, per, terms=("vol",))
    b = O.euler_dg_residual(m, U, GAMMA, 3, per)
    assert np.abs(vol).max() > 1e-3
    assert np.abs(b).max() <= 1e-12 * np.abs(vol).max()
    # a face left out (here one periodic face) leaves its cells unbalanced
    b2 = O.euler_dg_residual(m, U, GAMMA, 3, per[:-1])
    assert np.abs(b2).max() > 1e-3 * np.abs(vol).max()


@pytest.mark.parametrize("split", [False, True])
def test_conservation_on_a_periodic_mesh(split):
    n = 4
    m, per = _periodic_mesh(n, 0.15, seed=2, rotate=True, split=split)
    nc = len(m.celltypes)
    U = _state(np.random.default_rng(4), nc * 4).reshape(5, nc, 4)
    U[:, m.celltypes == O.CT["Tri3"], 3] = 0.0   # the unused slot of triangles
    b = O.euler_dg_residual(m, U, GAMMA, 3, per)
    face = O.euler_dg_residual(m, U, GAMMA, 3, per, terms=("face",))
    sums = b.sum(axis=(1, 2))
    assert np.abs(sums).max() <= 1e-12 * np.abs(face).max()


def test_periodic_pairing_and_mass():
    n = 3
    m, per = _periodic_mesh(n, 0.0)
    # West cells (x-index 0) pair with East cells (x-index n-1) of the same row
    for cm, sm, cp, sp, flip in per[:n]:
        assert cm % n == 0 and cp % n == n - 1 and cm // n == cp // n
        assert (sm, sp, flip) == (3, 1, 1)
    # the rows of a cell's Q1 mass matrix sum to area/4 (Σ_j φ_j = 1, ∫φ_i = area/4
    # on a parallelogram), so M⁻¹ 1 = (4/area) 1
    R = np.ones((5, n * n, 4))
    x = O.euler_dg_mass_inverse_apply(m, R, 3)
    area = (0.1 / n) ** 2
    assert np.allclose(x, 4.0 / area, rtol=1e-12)
