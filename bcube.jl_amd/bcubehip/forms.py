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
    def __mul__(self, o):
        return Mul(self, _wrap(o))

    def __rmul__(self, o):
        return Mul(_wrap(o), self)

    def __add__(self, o):
        return Add(self, _wrap(o))

    def __radd__(self, o):
        return Add(_wrap(o), self)


def _wrap(x):
    if isinstance(x, Expr):
        return x
    if isinstance(x, numbers.Real):
        return Const(float(x))
    raise TypeError(f"cannot use {type(x).__name__} in a form")


class Const(Expr):
    def __init__(self, v):
        self.v = v


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
    def __init__(self, role, space):
        self.role = role  # "trial" or "test"
        self.space = space


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


def grad(a):
    """∇ (gradient.jl:28)."""
    return Grad(a)


nabla = grad


def dot(a, b):
    """⋅ between two vector-valued operands."""
    return Dot(_wrap(a), _wrap(b))


class CellDomain:
    """CellDomain(mesh) (domain.jl:163-198)."""

    def __init__(self, mesh):
        self.mesh = mesh


class Measure:
    """Measure(domain, degree): Gauss-Legendre of that degree (measure.jl:30-55)."""

    def __init__(self, domain, degree: int):
        if not isinstance(domain, CellDomain):
            raise NotImplementedError("only CellDomain measures are supported")
        self.domain = domain
        self.degree = int(degree)


class Integral:
    def __init__(self, expr, measure):
        self.expr = expr
        self.measure = measure

    def __add__(self, o):
        return MultiIntegration([self]) + o

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
            if kind == "qp":
                raise NotImplementedError("stiffness coefficient must be constant or per cell")
            return "stiffness", kind, data, scale
    elif not dots:
        us = [f for f in rest if _is(f, Placeholder, "trial")]
        vs = [f for f in rest if _is(f, Placeholder, "test")]
        if len(us) == 1 and len(vs) == 1:
            kind, data, scale = _coef([f for f in rest if f is not us[0] and f is not vs[0]])
            if kind == "qp":
                raise NotImplementedError("mass coefficient must be constant or per cell")
            return "mass", kind, data, scale
    raise NotImplementedError("bilinear integrand is not one of: η ∇u⋅∇v, c u v")


def lower_linear(expr):
    """Recognise f v (source).  Returns (form, kind, data, scale)."""
    fs = _factors(expr, [])
    vs = [f for f in fs if _is(f, Placeholder, "test")]
    if len(vs) == 1 and not any(isinstance(f, (Dot, Grad, Add)) for f in fs):
        kind, data, scale = _coef([f for f in fs if f is not vs[0]])
        return "source", kind, data, scale
    raise NotImplementedError("linear integrand is not of the form f v")
