"""Host-side pieces of the DG transport path: recognising the form and the
nodal coordinates of a discontinuous space (no GPU needed)."""
import numpy as np
import pytest

import bcubehip as B
from bcubehip.forms import Placeholder, lower_dg_transport

C = np.array([1.0, 0.5, 0.25])
HOST = B.BcubeBackendHostOnly()


def _form(m, sign=-1.0, c_face=C, deg_face=7, swap=False):
    V = B.TrialFESpace(B.FunctionSpace("Lagrange", 3), m, "discontinuous")
    dΩ = B.Measure(B.CellDomain(m), 7)
    dΓ = B.Measure(B.InteriorFaceDomain(m), deg_face)
    nΓ = B.get_face_normals(dΓ)
    u = B.FEFunction(V, np.zeros((2, 10)))
    flux = B.UpwindFlux(c_face)
    a = (B.side_p(u), B.side_n(u)) if swap else (B.side_n(u), B.side_p(u))

    def l(v):
        vol = B.integrate(B.dot(C * u, B.grad(v)), dΩ)
        face = B.integrate(B.compose(flux, a[0], a[1], B.jump(v), B.side_n(nΓ)), dΓ)
        return vol - face if sign < 0 else vol + face

    return l(Placeholder("test", V)).terms(), u


def test_upwind_transport_form_is_recognised():
    m = B.hexa27_mesh(3, 3, 3, backend=HOST)
    items, u = _form(m)
    c, uu, deg = lower_dg_transport(items)
    assert np.array_equal(c, C) and uu is u and deg == 7


@pytest.mark.parametrize("kw", [dict(sign=1.0), dict(c_face=2 * C), dict(swap=True)])
def test_other_forms_are_not_taken_for_it(kw):
    m = B.hexa27_mesh(3, 3, 3, backend=HOST)
    items, _ = _form(m, **kw)
    assert lower_dg_transport(items) is None


def test_different_face_rule_is_refused():
    m = B.hexa27_mesh(3, 3, 3, backend=HOST)
    items, _ = _form(m, deg_face=5)
    with pytest.raises(NotImplementedError):
        lower_dg_transport(items)


def test_upwind_flux_value():
    f = B.UpwindFlux(C)
    n = np.array([0.0, 0.0, 1.0])
    assert f(2.0, 5.0, 3.0, n) == pytest.approx(0.25 * 2.0 * 3.0)
    assert f(2.0, 5.0, 3.0, -n) == pytest.approx(-0.25 * 5.0 * 3.0)


@pytest.mark.parametrize("kind", ["hexa8", "hexa27"])
def test_dg_dof_coordinates_hit_the_cell_nodes(kind):
    m = (B.hexa_mesh(3, 3, 2, backend=HOST) if kind == "hexa8"
         else B.hexa27_mesh(3, 3, 2, backend=HOST))
    B.sinusoidal_map(m, 0.03, seed=1)
    V = B.TrialFESpace(B.FunctionSpace("Lagrange", 2), m, "discontinuous")
    x = B.dg_dof_coordinates(V).reshape(m.ncells, 27, 3)
    X = m.nodes()
    c2n = m.c2n.reshape(m.ncells, -1)
    # Q2 uniform nodes at the corners coincide with the cell's vertices
    corner = {0: 0, 2: 1, 8: 2, 6: 3, 18: 4, 20: 5, 26: 6, 24: 7}
    for c in range(m.ncells):
        for q, k in corner.items():
            np.testing.assert_allclose(x[c, q], X[c2n[c, k]], atol=1e-14)
    if kind == "hexa27":  # and the Hexa27 nodes are the Q2 nodes themselves
        ref = B.mesh._H27_REF
        for c in range(m.ncells):
            for k in range(27):
                i1, i2, i3 = ref[k] + 1
                np.testing.assert_allclose(x[c, i1 + 3 * i2 + 9 * i3], X[c2n[c, k]], atol=1e-14)
