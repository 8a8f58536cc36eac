"""Numbering and pattern built by kernels (bh_build.hip), for meshes that
change: bit for bit the host DofHandler's ids (which equal the oracle's and
the reference's fixtures), on quads P1-P4, mixed triangle/quad meshes with
cells rotated so shared edges run both ways (the oriented_cell_side rule),
Quad9, hexahedra Q1-Q3 and the 64^3 Q2 hexa mesh of the headline benchmark."""
import time

import numpy as np
import pytest

import bcubehip as B

pytestmark = pytest.mark.gpu


def _rotated(m, seed):
    rng = np.random.default_rng(seed)
    cells = []
    for c in range(m.ncells):
        nodes = m.cell_nodes(c).tolist()
        k = int(rng.integers(len(nodes)))
        nodes = nodes[k:] + nodes[:k]
        cells.append(nodes[::-1] if rng.random() < 0.5 else nodes)
    names = {0: "Quad4", 2: "Tri3"}
    return B.Mesh(m.nodes(), [names[int(t)] for t in m.celltypes], cells)


def _cases():
    out = []
    for p in (1, 2, 3, 4):
        out.append((f"quad P{p}", lambda: B.rectangle_mesh(9, 7), p))
        out.append((f"mixed P{p}",
                    lambda p=p: B.mixed_tri_quad_mesh(10, block=2, p_split=0.5, seed=p)[0], p))
        out.append((f"mixed rotated P{p}",
                    lambda p=p: _rotated(B.mixed_tri_quad_mesh(10, block=2, p_split=0.5,
                                                               seed=p)[0], p), p))
    for p in (1, 2, 3):
        out.append((f"hexa Q{p}", lambda: B.hexa_mesh(5, 4, 4), p))
    out.append(("quad9 Q2", lambda: B.rectangle_mesh(6, 5, order=2), 2))
    return out


@pytest.mark.parametrize("name,make,p", _cases(), ids=[c[0] for c in _cases()])
@pytest.mark.parametrize("continuous", [True, False])
def test_device_numbering_equals_host(name, make, p, continuous):
    m = make()
    fs = B.FunctionSpace("Lagrange", p)
    host = B.DofHandler(m, fs, 1, continuous)
    dev = B.DofHandler(m, fs, 1, continuous, device=True)
    assert dev.ndofs_tot == host.ndofs_tot
    assert np.array_equal(dev.offset, host.offset)
    assert np.array_equal(dev.iglob, host.iglob)


def test_device_numbering_headline_mesh():
    m = B.hexa_mesh(65, 65, 65)
    fs = B.FunctionSpace("Lagrange", 2)
    t0 = time.perf_counter()
    host = B.DofHandler(m, fs)
    t1 = time.perf_counter()
    dev = B.DofHandler(m, fs, device=True)  # first call: context and module set-up
    t2 = time.perf_counter()
    dev = B.DofHandler(m, fs, device=True)
    t3 = time.perf_counter()
    print(f"64^3 Q2 numbering: host {t1 - t0:.3f} s, device {t3 - t2:.3f} s "
          f"(first call {t2 - t1:.3f} s)")
    assert dev.ndofs_tot == host.ndofs_tot == 129 ** 3
    assert np.array_equal(dev.iglob, host.iglob)


def _form(m, p, deg):
    dΩ = B.Measure(B.CellDomain(m), deg)
    kc = B.MeshCellData(np.random.default_rng(p).uniform(0.5, 2.0, m.ncells))
    return lambda u, v: (B.integrate(kc * B.dot(B.grad(u), B.grad(v)), dΩ)
                         + B.integrate(u * v, dΩ))


_PAT = [("quad P1", lambda: B.rectangle_mesh(9, 7), 1, 2),
        ("quad P3", lambda: B.rectangle_mesh(9, 7), 3, 6),
        ("mixed P2", lambda: B.mixed_tri_quad_mesh(10, block=2, p_split=0.5, seed=2)[0], 2, 4),
        ("mixed rotated P4", lambda: _rotated(B.mixed_tri_quad_mesh(10, block=2, p_split=0.5,
                                                                    seed=4)[0], 4), 4, 8),
        ("triangles P1", lambda: B.mixed_tri_quad_mesh(12, block=12, p_split=1.0, seed=1)[0], 1, 2),
        ("hexa Q1", lambda: B.hexa_mesh(6, 5, 5), 1, 2),
        ("hexa Q2", lambda: B.hexa_mesh(6, 5, 5), 2, 4)]


@pytest.mark.parametrize("name,make,p,deg", _PAT, ids=[c[0] for c in _PAT])
def test_device_pattern_equals_host(name, make, p, deg):
    """Same CSC structure, and the gather lists and offsets built on the device
    give the same values bit for bit (FP64 and FP32)."""
    m = make()
    U = B.TrialFESpace(B.FunctionSpace("Lagrange", p), m)
    V = B.TestFESpace(U)
    a = _form(m, p, deg)
    ph = B.BilinearPlan(a, U, V)
    pd = B.BilinearPlan(a, U, V, build="device")
    assert (pd.nrows, pd.nnz, pd.max_row) == (ph.nrows, ph.nnz, ph.max_row)
    ch, rh = ph.csc_structure()
    cd, rd = pd.csc_structure()
    assert np.array_equal(ch, cd) and np.array_equal(rh, rd)
    for mode in ("fp64", "fp32"):
        assert np.array_equal(pd.assemble_host(mode=mode), ph.assemble_host(mode=mode)), mode


def test_device_pattern_headline_mesh():
    m = B.hexa_mesh(65, 65, 65)
    U = B.TrialFESpace(B.FunctionSpace("Lagrange", 2), m)
    V = B.TestFESpace(U)
    a = _form(m, 2, 5)
    t0 = time.perf_counter()
    ph = B.BilinearPlan(a, U, V)
    t1 = time.perf_counter()
    pd = B.BilinearPlan(a, U, V, build="device")
    t2 = time.perf_counter()
    print(f"64^3 Q2 pattern: host {t1 - t0:.3f} s, device {t2 - t1:.3f} s, nnz {pd.nnz}")
    ch, rh = ph.csc_structure()
    cd, rd = pd.csc_structure()
    assert pd.nnz == 135005697
    assert np.array_equal(ch, cd) and np.array_equal(rh, rd)
    vh = ph.assemble_host(mode="fp32")
    assert np.array_equal(pd.assemble_host(mode="fp32"), vh)
