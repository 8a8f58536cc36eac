"""Host-side analysis of the one-pass structured Q2 kernel (bh_pattern_onepass
on host-only patterns, no GPU): the grid inferred from the dof map alone, the
row templates, and the spaces it must refuse."""
import numpy as np
import pytest

import bcubehip as B

HOST = B.BcubeBackendHostOnly()


def _plan(m, order=2, degree=5):
    U = B.TrialFESpace(B.FunctionSpace("Lagrange", order), m)
    dΩ = B.Measure(B.CellDomain(m), degree)
    return B.BilinearPlan(lambda u, v: B.integrate(B.dot(B.grad(u), B.grad(v)), dΩ)
                          + B.integrate(u * v, dΩ), U, B.TestFESpace(U))


@pytest.mark.parametrize("dims", [(1, 1, 1), (2, 3, 2), (5, 4, 3), (9, 9, 9), (1, 7, 2)])
def test_grid_inferred_from_dofs(dims):
    nx, ny, nz = dims
    info = _plan(B.hexa_mesh(nx + 1, ny + 1, nz + 1, backend=HOST)).onepass()
    assert info["ok"] and (info["nx"], info["ny"], info["nz"]) == dims, info
    # the layers per workgroup split nz so that the grid gives about one
    # workgroup per compute unit (256 on the host-only analysis)
    ntiles = -(-nx // 4) * -(-ny // 4)
    nzs = min(nz, max(1, -(-256 // ntiles)))
    assert info["layers_per_group"] == -(-nz // nzs)


def test_row_templates_are_few_and_shared():
    """Interior rows of the same lattice parity share a template: the count
    stops growing once every side has interior cells."""
    counts = [_plan(B.hexa_mesh(n + 1, n + 1, n + 1, backend=HOST)).onepass()["templates"]
              for n in (5, 7, 10)]
    assert counts[0] == counts[1] == counts[2] and counts[0] <= 343


def test_hexa27_grid_accepted():
    info = _plan(B.hexa27_mesh(4, 3, 5, backend=HOST)).onepass()
    assert info["ok"] and (info["nx"], info["ny"], info["nz"]) == (3, 2, 4)


def test_refusals_say_why():
    q1 = _plan(B.hexa_mesh(4, 4, 4, backend=HOST), order=1, degree=2).onepass()
    assert not q1["ok"] and "Q2" in q1["why"]
    quad = _plan(B.rectangle_mesh(5, 5, backend=HOST), order=2, degree=4).onepass()
    assert not quad["ok"]
    g = B.hexa_mesh(4, 4, 4, backend=HOST)
    c2n = np.asarray(g.c2n).reshape(-1, 8)[::-1]
    rev = B.Mesh(g.nodes(), ["Hexa8"] * g.ncells, list(c2n), HOST)
    info = _plan(rev).onepass()
    assert not info["ok"] and info["why"]
