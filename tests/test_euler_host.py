"""Host side of the COVO Euler DG residual: the periodic face pairing of the
product (BoundaryFaceDomain(mesh, PeriodicBCType)) against the oracle's
restatement of _compute_periodicity, rectangle_mesh's boundary names, and the
recognition of the benchmark's linear forms."""
import numpy as np
import pytest

import bcubehip as B
from bcubehip.euler import lower_euler_dg
from bcubehip.forms import Placeholder
from oracle import oracle as O

L_HALF = 0.05


def _om(m):
    return O.OMesh(m.dim, m.coords, m.celltypes, m.c2n_ptr, m.c2n)


@pytest.mark.parametrize("n", [3, 8])
def test_periodic_pairs_match_the_oracle(n):
    m = B.rectangle_mesh(n + 1, n + 1, xmin=-L_HALF, xmax=L_HALF, ymin=-L_HALF, ymax=L_HALF,
                         bnd_names=("West", "East", "South", "North"),
                         backend=B.BcubeBackendHostOnly())
    assert set(m.bc_nodes) == {"West", "East", "South", "North"}
    px = B.PeriodicBCType(B.Translation([-2 * L_HALF, 0.0]), ("East",), ("West",))
    py = B.PeriodicBCType(B.Translation([0.0, 2 * L_HALF]), ("South",), ("North",))
    for bc, shift in ((px, [-2 * L_HALF, 0.0]), (py, [0.0, 2 * L_HALF])):
        got = B.BoundaryFaceDomain(m, bc).periodic_faces()
        ref = O.periodic_faces(_om(m), m.bc_nodes, np.array(shift), bc.labels_master,
                               bc.labels_slave)
        assert np.array_equal(got, np.array(ref, dtype=np.int32))
        assert got.shape == (n, 5)


def test_lower_recognises_the_covo_forms():
    u = B.FEFunction(None, values=np.zeros(4))

    class _Mesh:
        pass

    mesh = _Mesh()
    dΩ = B.Measure(B.CellDomain(mesh), 3)
    dΓ = B.Measure(B.InteriorFaceDomain(mesh), 3)
    bc = B.PeriodicBCType(B.Translation([1.0, 0.0]), ("a",), ("b",))
    dΓp = B.Measure(B.BoundaryFaceDomain(mesh, bc), 3)
    v = tuple(Placeholder("test", None, field=i) for i in range(4))

    def fΩ(u, v):
        return B.compose(B.EulerFlux(1.4), u, tuple(map(B.grad, v)))

    def fΓ(u, v, n):
        return B.compose(B.RoeFlux(1.4), B.side_n(u), B.side_p(u), B.jump(v), B.side_n(n))

    n, npp = B.get_face_normals(dΓ), B.get_face_normals(dΓp)
    full = B.integrate(fΩ(u, v), dΩ) - B.integrate(fΓ(u, v, n), dΓ) - B.integrate(fΓ(u, v, npp), dΓp)
    g, uu, deg, per, terms, fc = lower_euler_dg(full.terms())
    assert (g, uu, deg, terms, fc) == (1.4, u, 3, 3, -1.0) and per == [dΓp.domain]
    _, _, _, per, terms, fc = lower_euler_dg(B.integrate(fΓ(u, v, n), dΓ).terms())
    assert (per, terms, fc) == ([], 2, 1.0)
    assert lower_euler_dg(B.integrate(fΩ(u, v), dΩ).terms())[4] == 1
    # other forms are not it
    assert lower_euler_dg(B.integrate(v[0] * 2.0, dΩ).terms()) is None
    with pytest.raises(NotImplementedError):   # periodic faces without the interior ones
        lower_euler_dg((-B.integrate(fΓ(u, v, npp), dΓp)).terms())
    # an integral given twice is refused, not dropped
    with pytest.raises(NotImplementedError):
        lower_euler_dg((B.integrate(fΩ(u, v), dΩ) + B.integrate(fΩ(u, v), dΩ)).terms())
    with pytest.raises(NotImplementedError):
        lower_euler_dg((B.integrate(fΓ(u, v, n), dΓ) + B.integrate(fΓ(u, v, n), dΓ)).terms())
    with pytest.raises(NotImplementedError):
        lower_euler_dg((B.integrate(fΓ(u, v, n), dΓ) + B.integrate(fΓ(u, v, npp), dΓp)
                        + B.integrate(fΓ(u, v, npp), dΓp)).terms())
