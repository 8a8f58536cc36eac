"""A small form language mirroring Bcube.jl's integrand DSL.

Julia                                Python (this package)
  ∫(η * ∇(u) ⋅ ∇(v))dΩ                 integrate(eta * dot(grad(u), grad(v)), dΩ)
  ∫(u * v)dΩ                           integrate(u * v, dΩ)
  ∫(f * v)dΩ                           integrate(f * v, dΩ)
  Measure(CellDomain(mesh), degree)    Measure(CellDomain(mesh), degree)
  MeshCellData(values)                 MeshCellData(values)
  PhysicalFunction(x -> ...)           PhysicalFunction(lambda x: ...)

A form is lowered to the closed set of integrands the HIP kernels implement
(stiffness, mass, source).  Anything else raises NotImplementedError: the
product path has no CPU fallback.  References: integration.jl:177-188
(Integrand/∫), measure.jl:30-55, meshdata.jl:23-36, cellfunction.jl:237-246,
gradient.jl:16-28.
"""
from __future__ import annotations

import numbers

import numpy as np


class Expr:
    __array_ufunc__ = None  # numpy arrays (a velocity c) defer to Expr.__rmul__

    def __mul__(self, o):
        return Mul(self, _wrap(o))

    def __rmul__(self, o):
        return Mul(_wrap(o), self)

    def __add__(self, o):
        return Add(self, _wrap(o))

    def __radd__(self, o):
        return Add(_wrap(o), self)

    def __neg__(self):
        return Mul(Const(-1.0), self)

    def __sub__(self, o):
        return Add(self, Mul(Const(-1.0), _wrap(o)))

    def __rsub__(self, o):
        return Add(_wrap(o), Mul(Const(-1.0), self))


def _wrap(x):
    if isinstance(x, Expr):
        return x
    if isinstance(x, numbers.Real):
        return Const(float(x))
    if isinstance(x, (tuple, list, np.ndarray)):
        return VecConst(x)
    raise TypeError(f"cannot use {type(x).__name__} in a form")


class Const(Expr):
    def __init__(self, v):
        self.v = v


class VecConst(Expr):
    """A constant vector, e.g. the velocity c in (c*u)."""

    def __init__(self, v):
        self.v = np.asarray(v, dtype=np.float64).reshape(-1)


class FEFunction(Expr):
    """FEFunction(U, values) (src/fespace/fefunction.jl): a discrete field of U.
    `values` has shape (ndofs,) or, for several fields evaluated together,
    (nfields, ndofs)."""

    def __init__(self, U, values=None, f=None):
        # FEFunction(U, mesh, f) (fefunction.jl): the nodal interpolant of a
        # PhysicalFunction on a continuous Lagrange space; FEFunction(U) is zero
        if f is not None:
            values = f
        self.space = getattr(U, "U", U)
        if values is None:
            values = np.zeros(self.space.get_ndofs(), dtype=np.float64)
        if isinstance(values, PhysicalFunction):
            from .dirichlet import dof_coordinates
            values = np.asarray(values.f(dof_coordinates(self.space)), dtype=np.float64)
        self.values = values

    def get_dof_values(self):
        """get_dof_values(u) (fefunction.jl)."""
        return self.values

    def set_dof_values_(self, values):
        """set_dof_values!(u, values) (fefunction.jl): copies into u's values."""
        values = np.asarray(values, dtype=np.float64)
        if values.shape != np.shape(self.values):
            raise ValueError("dof values have the wrong shape")
        self.values = values.copy()


class FaceNormal(Expr):
    """get_face_normals(dΓ) (domain.jl:677-690, FaceNormal cellfunction.jl:379-426)."""

    def __init__(self, measure):
        self.measure = measure


class Side(Expr):
    """side⁻(a) / side⁺(a) (cellfunction.jl:283-306)."""

    def __init__(self, which, a):
        self.which, self.a = which, a


class Jump(Expr):
    """jump(a) = side⁻(a) - side⁺(a) (cellfunction.jl:308)."""

    def __init__(self, a):
        self.a = a


class UpwindFlux:
    """The upwind transport flux (ui, uj, δv, n) -> (c⋅n > 0 ? (c⋅n)*ui : (c⋅n)*uj) * δv
    as a value for compose()."""

    def __init__(self, c):
        self.c = np.asarray(c, dtype=np.float64).reshape(-1)

    def __call__(self, ui, uj, dv, n):
        cn = float(np.dot(self.c, n))
        return (cn * ui if cn > 0 else cn * uj) * dv


class Compose(Expr):
    """f ∘ (a, b, ...) (Julia composition applied to a tuple of operators).  A
    tuple argument (the fields of a MultiFESpace, `map(∇, v)`) stays a tuple."""

    def __init__(self, f, args):
        self.f = f
        self.args = [tuple(_wrap(x) for x in a) if isinstance(a, tuple) else _wrap(a)
                     for a in args]


class EulerFlux:
    """_flux_Ω of the COVO benchmark (benchmark/covo.jl:44-58) as a value for
    compose(EulerFlux(γ), u, map(∇, v)): the Euler fluxes of (ρ, ρu, ρE, ϕ)
    against the test gradients."""

    def __init__(self, gamma):
        self.gamma = float(gamma)


class RoeFlux:
    """flux_roe of the COVO benchmark (benchmark/covo.jl:69-150) as a value for
    compose(RoeFlux(γ), side⁻(u), side⁺(u), jump(v), side⁻(n))."""

    def __init__(self, gamma):
        self.gamma = float(gamma)


def side_n(a):
    """side⁻ (cellfunction.jl:298-306); on a tuple, of each entry."""
    if isinstance(a, tuple):
        return tuple(side_n(x) for x in a)
    return Side("n", _wrap(a) if not isinstance(a, FaceNormal) else a)


def side_p(a):
    """side⁺ (cellfunction.jl:298-306); on a tuple, of each entry."""
    if isinstance(a, tuple):
        return tuple(side_p(x) for x in a)
    return Side("p", _wrap(a) if not isinstance(a, FaceNormal) else a)


side_minus, side_plus = side_n, side_p


def jump(a, n=None):
    """jump(a) = side⁻(a) - side⁺(a); jump(a, n) = side⁺(a) side⁺(n) + side⁻(a) side⁻(n)
    (cellfunction.jl:308-310)."""
    if n is None:
        if isinstance(a, tuple):   # jump(a::Tuple) = map(jump, a)
            return tuple(Jump(_wrap(x)) for x in a)
        return Jump(_wrap(a))
    return Add(Mul(side_p(a), side_p(n)), Mul(side_n(a), side_n(n)))


def avg(a):
    """avg(a) = (side⁺(a) + side⁻(a))/2, as bench_assemble.jl:31 defines it."""
    return Mul(Const(0.5), Add(side_p(a), side_n(a)))


def compose(f, *args):
    return Compose(f, args)


def get_face_normals(measure):
    return FaceNormal(measure)


class MeshCellData(Expr):
    """Per-cell values (meshdata.jl:23-36)."""

    def __init__(self, values):
        self.values = np.ascontiguousarray(values, dtype=np.float64)


class PhysicalFunction(Expr):
    """f(x) evaluated at physical quadrature points (cellfunction.jl:237-246).
    `f` takes an (n, dim) array of points and returns n values (vectorised)."""

    def __init__(self, f):
        self.f = f


class Placeholder(Expr):
    def __init__(self, role, space, field=None):
        self.role = role  # "trial" or "test"
        self.space = space
        self.field = field  # index inside a MultiFESpace, None for a single space


class Grad(Expr):
    def __init__(self, a):
        self.a = a


class Mul(Expr):
    def __init__(self, a, b):
        self.a, self.b = a, b


class Dot(Expr):
    def __init__(self, a, b):
        self.a, self.b = a, b


class Add(Expr):
    def __init__(self, a, b):
        self.a, self.b = a, b


class DDot(Expr):
    """⊡, the double contraction of two tensors (algebra.jl)."""

    def __init__(self, a, b):
        self.a, self.b = a, b


def grad(a):
    """∇ (gradient.jl:28)."""
    return Grad(a)


nabla = grad


def dot(a, b):
    """⋅ between two vector-valued operands."""
    return Dot(_wrap(a), _wrap(b))


def ddot(a, b):
    """⊡ between two tensor-valued operands, e.g. ∇(u) ⊡ ∇(v) on a vector space."""
    return DDot(_wrap(a), _wrap(b))


class CellDomain:
    """CellDomain(mesh) (domain.jl:163-198)."""

    def __init__(self, mesh):
        self.mesh = mesh


class InteriorFaceDomain:
    """InteriorFaceDomain(mesh) (domain.jl:186-198): faces with two cells."""

    def __init__(self, mesh):
        self.mesh = mesh


class Translation:
    """Translation(vec) (periodic transformations, x -> x + vec)."""

    def __init__(self, vec):
        self.vec = np.asarray(vec, dtype=np.float64).reshape(-1)

    def __call__(self, x):
        return np.asarray(x, dtype=np.float64) + self.vec


class PeriodicBCType:
    """PeriodicBCType(transformation, labels_master, labels_slave; tol)
    (boundary_condition.jl): slave faces are the images of master faces."""

    def __init__(self, transformation, labels_master, labels_slave, tol: float = 1e-9):
        self.transformation = transformation
        self.labels_master = tuple(labels_master)
        self.labels_slave = tuple(labels_slave)
        self.tol = float(tol)


class BoundaryFaceDomain:
    """BoundaryFaceDomain(mesh) (domain.jl:399-442): the faces with one cell, all
    boundary labels together; BoundaryFaceDomain(mesh, PeriodicBCType(...))
    (domain.jl:232-397): the slave faces, each with its master face's cell as
    side⁺."""

    def __init__(self, mesh, bc=None):
        # A label (the reference's BoundaryFaceDomain(mesh, label)) would select a
        # subset of the boundary faces; the face kernels integrate over all of
        # them, so labels are refused rather than silently widened.
        if bc is not None and not isinstance(bc, PeriodicBCType):
            raise NotImplementedError(
                "BoundaryFaceDomain takes no boundary label here: only None (every "
                "boundary face) or a PeriodicBCType")
        self.mesh = mesh
        self.bc = bc
        self._pairs = None

    def periodic_faces(self):
        """[cell⁻, side⁻, cell⁺, side⁺, flip] per slave face (int32, (n, 5))."""
        if self.bc is None:
            raise ValueError("not a periodic face domain")
        if self._pairs is None:
            from .euler import periodic_face_pairs
            self._pairs = periodic_face_pairs(self.mesh, self.bc)
        return self._pairs


class OrderDegree:
    """The Measure degree of a space of per-cell order: 2p on the cells of
    order p, at least `minimum` (no counterpart in the reference).  The
    library receives it as the degree -minimum."""

    def __init__(self, minimum: int = 1):
        if int(minimum) < 1:
            raise ValueError("minimum degree must be >= 1")
        self.minimum = int(minimum)


class Measure:
    """Measure(domain, degree): Gauss-Legendre of that degree (measure.jl:30-55).
    `degree` may be an OrderDegree on a space of per-cell order."""

    def __init__(self, domain, degree):
        if not isinstance(domain, (CellDomain, InteriorFaceDomain, BoundaryFaceDomain)):
            raise NotImplementedError(
                "only CellDomain, InteriorFaceDomain and BoundaryFaceDomain measures are supported")
        self.domain = domain
        self.degree = -degree.minimum if isinstance(degree, OrderDegree) else int(degree)


class Integral:
    def __init__(self, expr, measure):
        self.expr = expr
        self.measure = measure

    def __add__(self, o):
        return MultiIntegration([self]) + o

    def __sub__(self, o):
        return MultiIntegration([self]) - o

    def __neg__(self):
        return Integral(Mul(Const(-1.0), self.expr), self.measure)

    def terms(self):
        return [self]


class MultiIntegration:
    """Sum of integrals (integration.jl:199-314)."""

    def __init__(self, items):
        self.items = list(items)

    def __add__(self, o):
        if isinstance(o, Integral):
            return MultiIntegration(self.items + [o])
        if isinstance(o, MultiIntegration):
            return MultiIntegration(self.items + o.items)
        raise TypeError("can only add integrals")

    def __sub__(self, o):
        if isinstance(o, Integral):
            return self + (-o)
        if isinstance(o, MultiIntegration):
            return MultiIntegration(self.items + [-it for it in o.items])
        raise TypeError("can only subtract integrals")

    def terms(self):
        return list(self.items)


def integrate(expr, measure: Measure) -> Integral:
    """∫(expr)dΩ."""
    return Integral(_wrap(expr), measure)


# ---------------------------------------------------------------------------
# Lowering to library terms
# ---------------------------------------------------------------------------
def _factors(e, out):
    """Flatten a product tree into factors; Dot nodes stay as single factors."""
    if isinstance(e, Mul):
        _factors(e.a, out)
        _factors(e.b, out)
    else:
        out.append(e)
    return out


def _coef(factors):
    """Product of the coefficient factors -> (kind, array_or_fn, scale)."""
    scale, data, kind = 1.0, None, "const"
    for f in factors:
        if isinstance(f, Const):
            scale *= f.v
        elif isinstance(f, MeshCellData):
            if data is not None:
                raise NotImplementedError("product of two non-constant coefficients")
            data, kind = f.values, "cell"
        elif isinstance(f, PhysicalFunction):
            if data is not None:
                raise NotImplementedError("product of two non-constant coefficients")
            data, kind = f.f, "qp"
        elif isinstance(f, FEFunction):
            # materialised at the quadrature points on the device (bh_fe_eval_qp)
            if data is not None:
                raise NotImplementedError("product of two non-constant coefficients")
            if np.ndim(f.values) != 1:
                raise NotImplementedError("an FE function of several fields as a coefficient")
            data, kind = f, "fe"
        else:
            raise NotImplementedError(f"unsupported factor {type(f).__name__}")
    return kind, data, scale


def _is(x, cls, role=None):
    return isinstance(x, cls) and (role is None or x.role == role)


def lower_bilinear(expr):
    """Recognise η ∇u⋅∇v (stiffness) or c u v (mass).  Returns (form, kind, data, scale)."""
    fs = _factors(expr, [])
    dots = [f for f in fs if isinstance(f, Dot)]
    rest = [f for f in fs if not isinstance(f, Dot)]
    if len(dots) == 1:
        d = dots[0]
        # (η * ∇u) ⋅ ∇v : coefficients may sit inside either side of the dot
        left = _factors(d.a, [])
        right = _factors(d.b, [])
        inner = left + right
        gu = [f for f in inner if isinstance(f, Grad) and _is(f.a, Placeholder, "trial")]
        gv = [f for f in inner if isinstance(f, Grad) and _is(f.a, Placeholder, "test")]
        coefs = [f for f in inner if f not in gu and f not in gv] + rest
        if len(gu) == 1 and len(gv) == 1 and (gu[0] in left) != (gv[0] in left):
            kind, data, scale = _coef(coefs)
            if kind in ("qp", "fe"):
                raise NotImplementedError("stiffness coefficient must be constant or per cell")
            return "stiffness", kind, data, scale
        # u ⋅ v of scalar fields (Julia's dot of two numbers): c u v
        us = [f for f in inner if _is(f, Placeholder, "trial")]
        vs = [f for f in inner if _is(f, Placeholder, "test")]
        if (not gu and not gv and len(us) == 1 and len(vs) == 1 and (us[0] in left) != (vs[0] in left)
                and not any(isinstance(f, (Dot, Grad, Add, VecConst)) for f in inner + rest
                            if f is not us[0] and f is not vs[0])):
            kind, data, scale = _coef([f for f in inner + rest if f is not us[0] and f is not vs[0]])
            if kind in ("qp", "fe"):
                raise NotImplementedError("mass coefficient must be constant or per cell")
            return "mass", kind, data, scale
    elif not dots:
        us = [f for f in rest if _is(f, Placeholder, "trial")]
        vs = [f for f in rest if _is(f, Placeholder, "test")]
        if len(us) == 1 and len(vs) == 1:
            kind, data, scale = _coef([f for f in rest if f is not us[0] and f is not vs[0]])
            if kind in ("qp", "fe"):
                raise NotImplementedError("mass coefficient must be constant or per cell")
            return "mass", kind, data, scale
    raise NotImplementedError("bilinear integrand is not one of: η ∇u⋅∇v, c u v")


def lower_bilinear_terms(expr, dim):
    """lower_bilinear, plus the advection form α (c⋅∇u) v (c a constant
    vector; on a line also ∇(u)⋅v, c = 1) as one term per axis, form
    ("adv", a) with scale α c_a.  Returns a list."""
    fs = _factors(expr, [])
    dots = [f for f in fs if isinstance(f, Dot)]
    if len(dots) == 1:
        d = dots[0]
        left, right = _factors(d.a, []), _factors(d.b, [])
        rest = [f for f in fs if f is not d]
        for g, w in ((left, right), (right, left)):
            gu = [f for f in g if isinstance(f, Grad) and _is(f.a, Placeholder, "trial")]
            if len(gu) != 1:
                continue
            others = [f for f in g if f is not gu[0]] + w + rest
            vs = [f for f in others if _is(f, Placeholder, "test")]
            cs = [f for f in others if isinstance(f, VecConst)]
            coefs = [f for f in others if f not in vs and f not in cs]
            if len(vs) != 1 or len(cs) > 1 or any(isinstance(f, (Dot, Grad, Add, Placeholder))
                                                  for f in coefs):
                continue
            if cs:
                c = cs[0].v
            elif dim == 1:
                c = np.ones(1)
            else:
                continue
            kind, data, scale = _coef(coefs)
            if kind in ("qp", "fe"):
                raise NotImplementedError("advection coefficient must be constant or per cell")
            return [(("adv", a), kind, data, scale * float(ca)) for a, ca in enumerate(c)]
    return [lower_bilinear(expr)]


def _split_sum(e):
    """The product terms of an integrand: sums inside products distributed."""
    if isinstance(e, Add):
        return _split_sum(e.a) + _split_sum(e.b)
    if isinstance(e, Mul):
        return [Mul(x, y) for x in _split_sum(e.a) for y in _split_sum(e.b)]
    return [e]


def _placeholders(e, out):
    if isinstance(e, Placeholder):
        out.append(e)
    for k in ("a", "b"):
        if hasattr(e, k) and isinstance(getattr(e, k), Expr):
            _placeholders(getattr(e, k), out)
    return out


def _to_scalar_form(e, vector: bool):
    """A term over two vector spaces as the scalar form of each component:
    u⋅v -> u v and ∇u ⊡ ∇v -> ∇u⋅∇v, both component-diagonal.  On a vector
    space ∇u⋅∇v is the same Frobenius product as ∇u ⊡ ∇v: Bcube's ⋅ is
    LinearAlgebra.dot, which contracts all indices of two matrices
    (LazyOperators/LazyOperators.jl:26-27, algebra/algebra.jl:137).  Over scalar
    spaces ⊡ and ⋅ of two gradients agree, and u⋅v is u v."""
    if isinstance(e, Mul):
        return Mul(_to_scalar_form(e.a, vector), _to_scalar_form(e.b, vector))
    if isinstance(e, DDot):
        return Dot(e.a, e.b)
    if isinstance(e, Dot):
        pa, pb = _factors(e.a, []), _factors(e.b, [])
        if not any(isinstance(f, Grad) for f in pa + pb):
            return Mul(e.a, e.b)
    return e


def _grad_scalar_dot_vector(term, ncomp, dim):
    """α ∇(p)⋅v with p a scalar trial field and v a vector test field of dim
    components: per component a the scalar advection term α ∂p/∂x_a v_a."""
    fs = _factors(term, [])
    dots = [f for f in fs if isinstance(f, Dot)]
    if len(dots) == 1 and ncomp == dim:
        d = dots[0]
        rest = [f for f in fs if f is not d]
        for g, v in ((d.a, d.b), (d.b, d.a)):
            if (isinstance(g, Grad) and _is(g.a, Placeholder, "trial") and _is(v, Placeholder, "test")
                    and not any(isinstance(f, (Dot, Grad, Add, Placeholder, VecConst)) for f in rest)):
                kind, data, scale = _coef(rest)
                if kind in ("qp", "fe"):
                    raise NotImplementedError("coefficient must be constant or per cell")
                return [[(("adv", a), kind, data, scale)] for a in range(ncomp)]
    raise NotImplementedError("the only term coupling a scalar trial and a vector test space "
                              "lowered is ∇(p)⋅v")


def lower_multi_bilinear(items, sizes, dim=None):
    """Lower a bilinear form over several fields (assembler.jl:327-420).

    sizes[i]: components of space i, dim: the mesh dimension.  Returns
    {(test space i, trial space j): {(component ki, component kj): [(form,
    kind, data, scale), ...]}} for the pairs the form couples.  A pair of vector
    spaces gets the scalar form on each diagonal component pair (c u⋅v and
    η ∇u ⊡ ∇v); ∇(p)⋅v couples a scalar trial field p to each component a of a
    vector test field v by the advection term ∂p/∂x_a v_a.  Component pairs
    absent from the dict hold explicit zeros in the pattern."""
    out: dict[tuple[int, int], dict] = {}
    for it in items:
        for term in _split_sum(it.expr):
            ph = _placeholders(term, [])
            us = [p for p in ph if p.role == "trial"]
            vs = [p for p in ph if p.role == "test"]
            if len(us) != 1 or len(vs) != 1:
                raise NotImplementedError("each term must hold one trial and one test function")
            i, j = vs[0].field or 0, us[0].field or 0
            comp = out.setdefault((i, j), {})
            if sizes[i] > 1 and sizes[j] == 1:
                for a, forms in enumerate(_grad_scalar_dot_vector(term, sizes[i], dim)):
                    comp.setdefault((a, 0), []).extend(forms)
                continue
            if (sizes[i] > 1) != (sizes[j] > 1) or (sizes[i] > 1 and sizes[i] != sizes[j]):
                raise NotImplementedError("terms coupling a scalar and a vector space")
            lowered = lower_bilinear(_to_scalar_form(term, sizes[i] > 1))
            for k in range(sizes[i]):
                comp.setdefault((k, k), []).append(lowered)
    return out


def lower_multi_linear(items, sizes):
    """Lower l(v) over several fields: {(test space i, component k): [(form,
    kind, data, scale), ...]}; over a vector space the term is f⋅v with f a
    vector PhysicalFunction (component k of f.f's value per point) or a
    constant vector."""
    out: dict[tuple[int, int], list] = {}
    for it in items:
        for term in _split_sum(it.expr):
            vs = [p for p in _placeholders(term, []) if p.role == "test"]
            if len(vs) != 1:
                raise NotImplementedError("each term must hold one test function")
            i = vs[0].field or 0
            if sizes[i] == 1:
                t = _to_scalar_form(term, False)
                out.setdefault((i, 0), []).append(lower_linear(t))
                continue
            fs = _factors(term, [])
            dots = [f for f in fs if isinstance(f, Dot)]
            if len(dots) != 1:
                raise NotImplementedError("vector linear term must be f⋅v")
            d = dots[0]
            side = [d.a, d.b]
            if _is(side[0], Placeholder, "test"):
                side.reverse()
            f, v = side
            if not _is(v, Placeholder, "test"):
                raise NotImplementedError("vector linear term must be f⋅v")
            rest = [x for x in fs if x is not d]
            for k in range(sizes[i]):
                if isinstance(f, VecConst):
                    fk = Const(float(f.v[k]))
                elif isinstance(f, PhysicalFunction):
                    fk = PhysicalFunction(lambda x, f=f.f, k=k: np.asarray(f(x))[:, k])
                else:
                    raise NotImplementedError("f in f⋅v must be a PhysicalFunction or constant")
                e = v
                for x in rest + [fk]:
                    e = Mul(x, e)
                out.setdefault((i, k), []).append(lower_linear(e))
    return out


def lower_linear(expr):
    """Recognise f v (source).  Returns (form, kind, data, scale)."""
    fs = _factors(expr, [])
    vs = [f for f in fs if _is(f, Placeholder, "test")]
    if len(vs) == 1 and not any(isinstance(f, (Dot, Grad, Add)) for f in fs):
        kind, data, scale = _coef([f for f in fs if f is not vs[0]])
        return "source", kind, data, scale
    raise NotImplementedError("linear integrand is not of the form f v")


def lower_linear_terms(expr):
    """lower_linear, plus ((f c)⋅∇v) with c a constant vector: one term
    ∫ f ∂v/∂x_a per axis, scale c_a (form ("grad", a)).  Returns a list."""
    fs = _factors(expr, [])
    dots = [f for f in fs if isinstance(f, Dot)]
    if len(dots) != 1:
        return [lower_linear(expr)]
    d = dots[0]
    outer = [f for f in fs if f is not d]
    for g, w in ((_factors(d.a, []), _factors(d.b, [])), (_factors(d.b, []), _factors(d.a, []))):
        if len(g) == 1 and isinstance(g[0], Grad) and _is(g[0].a, Placeholder, "test"):
            cs = [f for f in w if isinstance(f, VecConst)]
            rest = [f for f in w if not isinstance(f, VecConst)] + outer
            if len(cs) == 1 and not any(isinstance(f, (Dot, Grad, Add, Placeholder, VecConst))
                                        for f in rest):
                kind, data, scale = _coef(rest)
                return [(("grad", a), kind, data, scale * float(c))
                        for a, c in enumerate(cs[0].v)]
    raise NotImplementedError("linear integrand is not of the form f v or (f c)⋅∇v")


def lower_dg_transport(items):
    """Recognise the upwind DG transport form

        ∫((c*u)⋅∇(v))dΩ - ∫(flux ∘ (side⁻(u), side⁺(u), jump(v), side⁻(nΓ)))dΓ

    (in either order).  Returns (c, u, degree) or None when the form is not it."""
    vol = face = None
    for it in items:
        if isinstance(it.measure.domain, CellDomain):
            vol = it if vol is None else False
        elif isinstance(it.measure.domain, InteriorFaceDomain):
            face = it if face is None else False
    if not vol or not face or len(items) != 2:
        return None
    fs = _factors(vol.expr, [])
    if len(fs) != 1 or not isinstance(fs[0], Dot):
        return None
    d = fs[0]
    sides = [_factors(d.a, []), _factors(d.b, [])]
    if isinstance(sides[0][0], Grad):
        sides.reverse()
    cu, g = sides
    if not (len(g) == 1 and isinstance(g[0], Grad) and _is(g[0].a, Placeholder, "test")):
        return None
    cs = [f for f in cu if isinstance(f, VecConst)]
    us = [f for f in cu if isinstance(f, FEFunction)]
    if len(cs) != 1 or len(us) != 1 or len(cu) != 2:
        return None
    c, u = cs[0].v, us[0]
    ff = _factors(face.expr, [])
    comp = [f for f in ff if isinstance(f, Compose)]
    sign = 1.0
    for f in ff:
        if isinstance(f, Const):
            sign *= f.v
    if len(comp) != 1 or len(ff) != 1 + sum(isinstance(f, Const) for f in ff) or sign != -1.0:
        return None
    cp = comp[0]
    if not isinstance(cp.f, UpwindFlux) or not np.array_equal(cp.f.c, c) or len(cp.args) != 4:
        return None
    a0, a1, a2, a3 = cp.args
    ok = (isinstance(a0, Side) and a0.which == "n" and a0.a is u and
          isinstance(a1, Side) and a1.which == "p" and a1.a is u and
          isinstance(a2, Jump) and _is(a2.a, Placeholder, "test") and
          isinstance(a3, Side) and a3.which == "n" and isinstance(a3.a, FaceNormal))
    if not ok or c.size != 3:
        return None
    if vol.measure.degree != face.measure.degree:
        raise NotImplementedError("the DG transport kernels use one Gauss rule on cells and faces")
    if vol.measure.domain.mesh is not face.measure.domain.mesh:
        raise NotImplementedError("cell and face measures must be on one mesh")
    return c, u, vol.measure.degree


# ---------------------------------------------------------------------------
# Face integrands of the interior-penalty kind, lowered symbolically
# ---------------------------------------------------------------------------
# A face integrand is expanded into a sum of products of atoms: the test
# function's value "v{s}" or gradient "gv{s}" on side s (0 = side⁻, the cell
# that created the face; 1 = side⁺), the same for the trial function ("u",
# "gu"), the outward normals "n{s}", and a PhysicalFunction "g:{id}".  A
# product keeps its scalar atoms (dot products "d:x|y" among them) and at most
# one vector atom.  The normals of the two sides of a straight face are
# opposite, n1 = -n0, so every normal is σ_s n_0 with σ_0 = 1, σ_1 = -1.
_SIGMA = {0: 1.0, 1: -1.0}


class _Sym:
    """Σ coef · Π scalars · vector; terms[(scalars (sorted tuple), vector or None)] = coef."""

    def __init__(self, terms=None):
        self.terms = dict(terms or {})

    @staticmethod
    def atom(name, vector=False):
        return _Sym({((), name) if vector else ((name,), None): 1.0})

    def rank(self):
        ranks = {vec is not None for (_, vec) in self.terms}
        if len(ranks) > 1:
            raise NotImplementedError("a face integrand adds a scalar and a vector")
        return ranks.pop() if ranks else False

    def add(self, o):
        t = dict(self.terms)
        for k, c in o.terms.items():
            t[k] = t.get(k, 0.0) + c
        out = _Sym(t)
        out.rank()
        return out

    def mul(self, o):
        t = {}
        for (sa, va), ca in self.terms.items():
            for (sb, vb), cb in o.terms.items():
                if va is not None and vb is not None:
                    raise NotImplementedError("product of two vectors in a face integrand "
                                              "(use dot)")
                k = (tuple(sorted(sa + sb)), va if va is not None else vb)
                t[k] = t.get(k, 0.0) + ca * cb
        return _Sym(t)

    def dot(self, o):
        t = {}
        for (sa, va), ca in self.terms.items():
            for (sb, vb), cb in o.terms.items():
                if va is None or vb is None:
                    raise NotImplementedError("dot of a scalar in a face integrand")
                d = "d:" + "|".join(sorted((va, vb)))
                k = (tuple(sorted(sa + sb + (d,))), None)
                t[k] = t.get(k, 0.0) + ca * cb
        return _Sym(t)


def _face_sym(e, side, gfuns):
    """Expand a face integrand; `side` is the side an enclosing side⁻/side⁺ picked."""
    if isinstance(e, Const):
        return _Sym({((), None): e.v})
    if isinstance(e, Add):
        return _face_sym(e.a, side, gfuns).add(_face_sym(e.b, side, gfuns))
    if isinstance(e, Mul):
        return _face_sym(e.a, side, gfuns).mul(_face_sym(e.b, side, gfuns))
    if isinstance(e, Dot):
        return _face_sym(e.a, side, gfuns).dot(_face_sym(e.b, side, gfuns))
    if isinstance(e, Side):
        return _face_sym(e.a, 0 if e.which == "n" else 1, gfuns)
    if isinstance(e, Jump):
        return _face_sym(e.a, 0, gfuns).add(_face_sym(Mul(Const(-1.0), e.a), 1, gfuns))
    if isinstance(e, Compose):
        if isinstance(e.f, UpwindFlux):
            raise NotImplementedError("the upwind flux is not an interior-penalty integrand")
        return _face_sym(_wrap(e.f(*e.args)), side, gfuns)
    if isinstance(e, PhysicalFunction):
        gfuns.setdefault(id(e), e)
        return _Sym.atom(f"g:{id(e)}")
    if side is None:
        raise NotImplementedError("a face integrand needs side⁻ / side⁺ on every function")
    if isinstance(e, Placeholder):
        return _Sym.atom(("v" if e.role == "test" else "u") + str(side))
    if isinstance(e, Grad) and isinstance(e.a, Placeholder):
        return _Sym.atom(("gv" if e.a.role == "test" else "gu") + str(side), vector=True)
    if isinstance(e, FaceNormal):
        return _Sym.atom("n" + str(side), vector=True)
    raise NotImplementedError(f"unsupported operand {type(e).__name__} in a face integrand")


def _face_monomials(e, boundary):
    """(coef, test kind, test side, trial kind or None, trial side, g) per product
    term; kinds "v" (value) or "dn" (n_0⋅∇), every other normal folded into coef."""
    gfuns = {}
    sym = _face_sym(e, 0 if boundary else None, gfuns)
    if sym.rank():
        raise NotImplementedError("a face integrand must be a scalar")
    out = []
    for (scalars, _), c in sym.terms.items():
        if c == 0.0:
            continue
        test = trial = g = None
        for a in scalars:
            if a.startswith("g:"):
                if g is not None:
                    raise NotImplementedError("product of two functions in a face integrand")
                g = gfuns[int(a[2:])]
                continue
            if a.startswith("d:"):
                x, y = a[2:].split("|")
                if x[0] == "n" and y[0] == "n":
                    c *= _SIGMA[int(x[1:])] * _SIGMA[int(y[1:])]
                    continue
                nrm, grd = (x, y) if x[0] == "n" else (y, x)
                if nrm[0] != "n":
                    raise NotImplementedError("∇u⋅∇v on faces")
                c *= _SIGMA[int(nrm[1:])]
                kind, s = "dn", int(grd[-1])
                role = grd[1]
            else:
                kind, s, role = "v", int(a[-1]), a[0]
            slot = (kind, s)
            if role == "v":
                if test is not None:
                    raise NotImplementedError("two test functions in one face term")
                test = slot
            else:
                if trial is not None:
                    raise NotImplementedError("two trial functions in one face term")
                trial = slot
        if test is None:
            raise NotImplementedError("a face term without the test function")
        out.append((c, test, trial, g))
    return out


def lower_face_bilinear(expr, boundary: bool):
    """Coefficients of a bilinear face integrand in the monomials the face
    kernels evaluate: per (test side t, trial side r) the 4 values
    (A, B, C, E) of  A (n_t⋅n_r) v_t u_r + B v_t (n_t⋅∇u_r) + C (n_r⋅∇v_t) u_r
    + E v_t u_r,  tr = t*2 + r; 16 values on interior faces, 4 on boundary
    faces (t = r = 0).  Integrands are sums of products of side⁻/side⁺ values
    and gradients of u and v, normals and constants (jump, avg and Python
    functions applied with compose expand to these)."""
    coef = np.zeros(16 if not boundary else 4)
    for c, test, trial, g in _face_monomials(expr, boundary):
        if trial is None or g is not None:
            raise NotImplementedError("a bilinear face term needs u and v and no function of x")
        (tk, t), (rk, r) = test, trial
        base = 4 * (2 * t + r) if not boundary else 0
        if tk == "v" and rk == "v":
            coef[base + 3] += c
        elif tk == "v" and rk == "dn":
            coef[base + 1] += c * _SIGMA[t]      # n_0 = σ_t n_t
        elif tk == "dn" and rk == "v":
            coef[base + 2] += c * _SIGMA[r]      # n_0 = σ_r n_r
        else:
            raise NotImplementedError("(n⋅∇v)(n⋅∇u) face terms")
    return coef


def lower_face_linear(expr):
    """A boundary linear face integrand as [(c0, c1, g)]: Σ (c0 v + c1 n⋅∇v) g,
    one entry per PhysicalFunction g (None: g = 1)."""
    groups = {}
    for c, test, trial, g in _face_monomials(expr, True):
        if trial is not None:
            raise NotImplementedError("a linear face term holds no trial function")
        kind, _ = test
        c0, c1, _g = groups.get(id(g), (0.0, 0.0, g))
        if kind == "v":
            c0 += c
        else:
            c1 += c
        groups[id(g)] = (c0, c1, g)
    return list(groups.values())
