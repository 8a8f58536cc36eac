"""Launch-shape settings on a host-only pattern: only the pattern's candidate
gather shapes are accepted (no GPU needed)."""
import pytest

import bcubehip as B

HOST = B.BcubeBackendHostOnly()


def _plan(order=2):
    m = B.hexa_mesh(4, 4, 4, backend=HOST)
    U = B.TrialFESpace(B.FunctionSpace("Lagrange", order), m)
    dΩ = B.Measure(B.CellDomain(m), 4)
    return B.BilinearPlan(lambda u, v: B.integrate(B.dot(B.grad(u), B.grad(v)), dΩ), U,
                          B.TestFESpace(U), "fp32")


def test_untuned_then_fixed_shapes():
    p = _plan()
    assert p.tuning() is None
    p.set_tuning(12, 32, 2)
    t = p.tuning()
    assert (t["elem_waves"], t["gather_lanes"], t["gather_rows"]) == (12, 32, 2)
    p.set_tuning()  # back to the defaults, still marked as chosen
    assert p.tuning()["gather_lanes"] == 0


@pytest.mark.parametrize("shape", [(7, 0, 0), (8, 5, 1), (8, 32, 3), (8, 4, 1)])
def test_rejects_non_candidates(shape):
    # Q2 hexa rows: 27-entry element rows take 32 lanes only
    p = _plan()
    with pytest.raises(B.BcubeHipError):
        p.set_tuning(*shape)
