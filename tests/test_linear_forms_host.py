"""Host-side lowering of the linear integrands added for the reference's
known answers (test_assembler.jl:40-68, 241-269): FE functions as
coefficients and (f c)⋅∇v, one ∫ f ∂v/∂x_a term per axis."""
import numpy as np
import pytest

import bcubehip as B
from bcubehip.assembler import _integrals
from bcubehip.forms import Placeholder, lower_bilinear, lower_linear_terms


def _space(p=1, cont=True):
    m = B.one_cell_mesh(":quad", backend=B.BcubeBackendHostOnly())
    return m, B.TrialFESpace(B.FunctionSpace("Lagrange", p), m, isContinuous=cont)


def _lower(l, U):
    return [t for it in _integrals(l(Placeholder("test", U))) for t in lower_linear_terms(it.expr)]


def test_transport_form_lowers_per_axis():
    m, U = _space(cont=False)
    u = B.FEFunction(U, np.ones(4))
    dΩ = B.Measure(B.CellDomain(m), 3)
    c = np.array([0.5, 1.0])
    for l in (lambda v: B.integrate(B.dot(c * u, B.grad(v)), dΩ),
              lambda v: B.integrate(2.0 * B.dot(B.grad(v), u * c), dΩ)):
        t = _lower(l, U)
        assert [x[0] for x in t] == [("grad", 0), ("grad", 1)]
        assert all(x[1] == "fe" and x[2] is u for x in t)
        s = 2.0 if t[1][3] == 2.0 else 1.0
        assert [x[3] for x in t] == [0.5 * s, 1.0 * s]


def test_fe_function_coefficient_and_refusals():
    m, U = _space(p=2)
    u = B.FEFunction(U, np.arange(9.0))
    dΩ = B.Measure(B.CellDomain(m), 4)
    assert _lower(lambda v: B.integrate(u * v, dΩ), U) == [("source", "fe", u, 1.0)]
    # two non-constant coefficients, and FE-function coefficients in bilinear forms
    with pytest.raises(NotImplementedError):
        _lower(lambda v: B.integrate(u * B.MeshCellData(np.ones(1)) * v, dΩ), U)
    ex = B.integrate(u * Placeholder("trial", U) * Placeholder("test", U), dΩ)
    with pytest.raises(NotImplementedError):
        lower_bilinear(_integrals(ex)[0].expr)
    with pytest.raises(NotImplementedError):
        _lower(lambda v: B.integrate(B.dot(B.grad(v), B.grad(v)), dΩ), U)


def test_scalar_dot_is_mass():
    """u ⋅ v of scalar fields (test_assembler.jl:233-239) is the mass form."""
    m, U = _space(p=2)
    dΩ = B.Measure(B.CellDomain(m), 2)
    ex = B.integrate(B.dot(Placeholder("trial", U), 3.0 * Placeholder("test", U)), dΩ)
    assert lower_bilinear(_integrals(ex)[0].expr) == ("mass", "const", None, 3.0)
    ex = B.integrate(B.dot(B.grad(Placeholder("trial", U)), Placeholder("test", U)), dΩ)
    with pytest.raises(NotImplementedError):
        lower_bilinear(_integrals(ex)[0].expr)


def test_advection_lowers_per_axis():
    from bcubehip.forms import lower_bilinear_terms
    m, U = _space(p=1)
    dΩ = B.Measure(B.CellDomain(m), 2)
    c = np.array([0.5, -2.0])
    u, v = Placeholder("trial", U), Placeholder("test", U)
    for ex in (B.integrate(B.dot(c, B.grad(u)) * v, dΩ),
               B.integrate(v * B.dot(B.grad(u), c), dΩ)):
        t = lower_bilinear_terms(_integrals(ex)[0].expr, 2)
        assert [x[0] for x in t] == [("adv", 0), ("adv", 1)]
        assert [x[3] for x in t] == [0.5, -2.0]
    # on a line, ∇(u)⋅v is ∂u/∂x v (test_assembler.jl:214-231)
    ex = B.integrate(B.dot(B.grad(u), v), dΩ)
    assert lower_bilinear_terms(_integrals(ex)[0].expr, 1) == [(("adv", 0), "const", None, 1.0)]
    with pytest.raises(NotImplementedError):
        lower_bilinear_terms(_integrals(ex)[0].expr, 2)
