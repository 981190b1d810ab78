"""Workload definitions: the benchmark problems as neutral specs, and their
Formulation<T> on the engine.

  poisson2d      -Laplace(u) = 1 on the unit square, P_k triangles, Dirichlet
                 u = 0 on the boundary (64x64 cells, P1, Gauss2 is the
                 2-D Poisson workload)
  helmholtz3d    -Laplace(u) - k^2 u = f on the unit cube, P_k tetrahedra,
                 Robin du/dn - i k u = 0 on the six faces, Gaussian source
                 f = exp(-|x - c|^2 / (2 s^2)) (50^3 cells x 6 tets, P2,
                 k = 10 pi, s = 0.05 is the 3-D Helmholtz workload)
  highorder3d    mass + stiffness on order-4 tetrahedra, Gauss8

Terms follow the reference's sign convention: the sum of the integral()
terms is zero and the assembled system is A x = -b.
"""
from dataclasses import dataclass, field
import math

from . import (Domain, Field, Formulation, dof, exp, grad, set_current_model, tf, unit_cube, unit_square, x, y,
               z)


@dataclass
class Spec:
    name: str
    model: object
    order: int
    field_domain: list
    dirichlet: list = field(default_factory=list)  # [(entities, value)]
    terms: list = field(default_factory=list)
    symmetric: bool = False
    frequency: float = 0.0
    dofs_sort: str = "default"
    scalar: str = "f64"
    basis: str = "lagrange"  # or "hierarchical" (HierarchicalH1 on a first-order mesh)

    def element_dofs(self):
        """sum over assembled elements of dofs per element (element-DOFs)."""
        ents = set()
        for t in self.terms:
            ents.update(t["domain"])
        total = 0
        for b in self.model.blocks():
            if (b["dim"], b["tag"]) in ents and b["dim"] > 0:
                npe = b["conn"].shape[1] if self.basis == "lagrange" else _hier_count(b["dim"], self.order)
                total += b["conn"].shape[0] * npe
        return total


    def scatter_map_bytes(self):
        """bytes of the reference-format element maps that the matrix scatter
        reads, over the elements of bilinear terms: element->DOF (4 B per local
        DOF) and element->nonzero (4 B per local pair), E (4 N + 4 N^2)."""
        ents = set()
        for t in self.terms:
            if t["kind"] == "bilinear":
                ents.update(t["domain"])
        total = 0
        for b in self.model.blocks():
            if (b["dim"], b["tag"]) in ents and b["dim"] > 0:
                npe = b["conn"].shape[1] if self.basis == "lagrange" else _hier_count(b["dim"], self.order)
                total += b["conn"].shape[0] * (4 * npe + 4 * npe * npe)
        return total


def _hier_count(dim, p):
    return p + 1 if dim == 1 else ((p + 1) * (p + 2) // 2 if dim == 2 else (p + 1) * (p + 2) * (p + 3) // 6)


def poisson2d(n=64, order=1, jitter=0.2, seed=1234, gauss=None, basis="lagrange"):
    m = unit_square(n, 1 if basis == "hierarchical" else order, jitter, seed)
    g = 2 * order if gauss is None else gauss
    omega = [(2, 1)]
    gamma = [(1, 1), (1, 2), (1, 3), (1, 4)]
    return Spec(
        name=f"poisson2d_n{n}_p{order}" + ("_h" if basis == "hierarchical" else ""), model=m, order=order,
        field_domain=omega,
        dirichlet=[(gamma, 0.0)],
        terms=[
            dict(kind="bilinear", domain=omega, gauss=g, form=1),
            dict(kind="linear", domain=omega, gauss=g, source=-1.0),
        ],
        scalar="f64", basis=basis)


def gaussian_source(center=(0.5, 0.5, 0.5), sigma=0.05):
    cx, cy, cz = center
    r2 = (x() - cx) * (x() - cx) + (y() - cy) * (y() - cy) + (z() - cz) * (z() - cz)
    return exp(-(r2 / (2.0 * sigma * sigma)))


def helmholtz3d(n=50, order=2, k=2 * math.pi * 5, jitter=0.15, seed=4321, center=(0.5, 0.5, 0.5), sigma=0.05,
                scalar="c128", gauss=None, basis="lagrange"):
    m = unit_cube(n, 1 if basis == "hierarchical" else order, jitter, seed)
    g = 2 * order if gauss is None else gauss
    omega = [(3, 1)]
    gamma = [(2, s) for s in range(1, 7)]
    f = gaussian_source(center, sigma)
    return Spec(
        name=f"helmholtz3d_n{n}_p{order}" + ("_h" if basis == "hierarchical" else ""), model=m, order=order,
        field_domain=omega,
        terms=[
            dict(kind="bilinear", domain=omega, gauss=g, form=1),
            dict(kind="bilinear", domain=omega, gauss=g, form=0, coef=-k * k),
            dict(kind="bilinear", domain=gamma, gauss=g, form=0, coef=-1j * k),
            dict(kind="linear", domain=omega, gauss=g, source=-f),
        ],
        scalar=scalar, basis=basis)


# term tags of helmholtz3d (order of its integral() calls)
HELMHOLTZ_MASS, HELMHOLTZ_ROBIN, HELMHOLTZ_SOURCE = 1, 2, 3


def helmholtz_params(k, center=(0.5, 0.5, 0.5), sigma=0.05):
    """Term coefficients of one Helmholtz request on a helmholtz3d
    formulation: wavenumber k and Gaussian source position."""
    from . import ScalarFunction
    return [(HELMHOLTZ_MASS, ScalarFunction(-k * k)), (HELMHOLTZ_ROBIN, ScalarFunction(-1j * k)),
            (HELMHOLTZ_SOURCE, -gaussian_source(center, sigma))]


def highorder3d(n=32, order=4, jitter=0.15, seed=99, gauss=8, scalar="f64"):
    m = unit_cube(n, order, jitter, seed)
    omega = [(3, 1)]
    return Spec(
        name=f"highorder3d_n{n}_p{order}", model=m, order=order, field_domain=omega,
        terms=[
            dict(kind="bilinear", domain=omega, gauss=gauss, form=1),
            dict(kind="bilinear", domain=omega, gauss=gauss, form=0),
        ],
        scalar=scalar)


def _domain(ents):
    d = Domain()
    for dim, tag in ents:
        d = d | Domain(dim, tag)
    return d


def build(spec, scalar=None):
    """Formulation of a spec on the current device, before pre()."""
    set_current_model(spec.model)
    F = Formulation(spec.name, scalar or spec.scalar)
    u = Field("u", _domain(spec.field_domain), spec.order, spec.basis)
    for ents, value in spec.dirichlet:
        u.add_constraint(_domain(ents), value)
    F._field = u  # keep the field alive with the formulation
    for t in spec.terms:
        dom = _domain(t["domain"])
        gs = f"Gauss{t['gauss']}"
        herm = t.get("hermitian", True)
        if t["kind"] == "bilinear":
            if t.get("dt", 0) == 1:
                from . import dt_dof as dd
            elif t.get("dt", 0) == 2:
                from . import dt2_dof as dd
            else:
                dd = dof
            lhs, rhs = dd(u), tf(u)
            if t["form"] == 1:
                lhs, rhs = grad(lhs), grad(rhs)
            c = t.get("coef", 1.0)
            if not (isinstance(c, (int, float, complex)) and c == 1.0):
                lhs = c * lhs
            tc = t.get("tf_coef", 1.0)
            if tc != 1.0:
                rhs = tc * rhs
            F.integral(lhs, rhs, dom, gs, herm)
        else:
            rhs = tf(u)
            tc = t.get("tf_coef", 1.0)
            if tc != 1.0:
                rhs = tc * rhs
            F.integral(t["source"], rhs, dom, gs, herm, t.get("rhs_idx", 0))
    F.set_symmetric(spec.symmetric)
    F.set_angular_frequency(spec.frequency)
    return F
