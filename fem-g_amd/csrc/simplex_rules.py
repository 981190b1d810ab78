"""Derive the symmetric simplex quadrature tables of refelem.h (and the
oracle's copy): fully symmetric orbits, positive weights, interior points,
exact for every monomial of the stated degree on the gmsh reference triangle
(area 1/2) and tetrahedron (volume 1/6).

The orbit structures are those of the published rules (Strang & Fix and
Dunavant 1985 for triangles; Keast 1986 and Walkington 2000 for tetrahedra).
The starting values are rounded published values; the moment equations are
then solved to machine precision.  Output: C rows (u, v, w, weight), points
ordered by decreasing weight, then coordinates.

Run: python fem-g_amd/csrc/simplex_rules.py > rules.inc   (numpy + scipy)
"""
import itertools
import math

import numpy as np
from scipy.optimize import least_squares


def orbit_points(dim, orbits, p):
    """barycentric points and weights of the orbits with parameters p"""
    pts, w, k = [], [], 0
    for o in orbits:
        if o in ("S3", "S4"):
            pts.append((1.0 / (dim + 1),) * (dim + 1))
            w.append(p[k])
            k += 1
            continue
        if o == "S21":
            a, wt = p[k], p[k + 1]
            base = (a, a, 1 - 2 * a)
            k += 2
        elif o == "S111":
            a, b, wt = p[k], p[k + 1], p[k + 2]
            base = (a, b, 1 - a - b)
            k += 3
        elif o == "S31":
            a, wt = p[k], p[k + 1]
            base = (a, a, a, 1 - 3 * a)
            k += 2
        elif o == "S22":
            a, wt = p[k], p[k + 1]
            base = (a, a, 0.5 - a, 0.5 - a)
            k += 2
        elif o == "S211":
            a, b, wt = p[k], p[k + 1], p[k + 2]
            base = (a, a, b, 1 - 2 * a - b)
            k += 3
        for t in sorted(set(itertools.permutations(base))):
            pts.append(t)
            w.append(wt)
    return np.array(pts), np.array(w)


def moments_residual(dim, orbits, deg, p):
    L, w = orbit_points(dim, orbits, p)
    x = L[:, 1:]  # (u, v[, w]) = barycentric coordinates 1..dim
    r = []
    for a in itertools.product(range(deg + 1), repeat=dim):
        if sum(a) > deg:
            continue
        exact = math.prod(math.factorial(ai) for ai in a) / math.factorial(sum(a) + dim)
        r.append(np.dot(w, np.prod(x ** np.array(a), axis=1)) - exact)
    return np.array(r)


def solve(dim, orbits, deg, p0):
    p = least_squares(lambda q: moments_residual(dim, orbits, deg, q), p0, xtol=1e-15, ftol=1e-15,
                      gtol=1e-15, max_nfev=20000).x
    for _ in range(8):  # Gauss-Newton polish, central-difference Jacobian
        r = moments_residual(dim, orbits, deg, p)
        J = np.empty((len(r), len(p)))
        for k in range(len(p)):
            h = 1e-7 * max(1.0, abs(p[k]))
            e = np.zeros(len(p))
            e[k] = h
            J[:, k] = (moments_residual(dim, orbits, deg, p + e) - moments_residual(dim, orbits, deg, p - e)) / (2 * h)
        p = p - np.linalg.lstsq(J, r, rcond=None)[0]
    L, w = orbit_points(dim, orbits, p)
    res = np.abs(moments_residual(dim, orbits, deg, p)).max()
    assert res < 1e-15 and (w > 0).all() and (L > 0).all(), (dim, deg, res)
    idx = sorted(range(len(w)), key=lambda i: (-round(w[i], 14), tuple(np.round(L[i], 14))))
    return L[idx][:, 1:], w[idx]


V = 1.0 / 6.0
RULES = [  # (name, dim, degree, orbits, starting parameters: orbit coordinates, weights)
    ("kTri1", 2, 1, ["S3"], [0.5]),
    ("kTri2", 2, 2, ["S21"], [1 / 6, 1 / 6]),
    ("kTri4", 2, 4, ["S21", "S21"], [0.445948490915965, 0.111690794839005, 0.091576213509771, 0.054975871827661]),
    ("kTri5", 2, 5, ["S3", "S21", "S21"],
     [0.1125, 0.470142064105115, 0.066197076394253, 0.101286507323456, 0.062969590272413]),
    ("kTri6", 2, 6, ["S21", "S21", "S111"],
     [0.249286745170910, 0.058393137863189, 0.063089014491502, 0.025422453185103, 0.310352451033785,
      0.053145049844816, 0.041425537809187]),
    ("kTri8", 2, 8, ["S3", "S21", "S21", "S21", "S111"],
     [0.0721578038388935, 0.459292588292723, 0.0475458171336425, 0.170569307751760, 0.051608685267359,
      0.050547228317031, 0.016229248811599, 0.263112829634638, 0.008394777409958, 0.0136151570872175]),
    ("kTet1", 3, 1, ["S4"], [V]),
    ("kTet2", 3, 2, ["S31"], [0.1381966011250105, V / 4]),
    ("kTet5", 3, 5, ["S31", "S31", "S22"],
     [0.0927352503108912, 0.01224884051939366, 0.3108859192633006, 0.01878132095300264, 0.4544962958743504,
      0.007091003462846911]),
    ("kTet6", 3, 6, ["S31", "S31", "S31", "S211"],
     [0.2146028712591520, 0.0399227502581679 * V, 0.0406739585346113, 0.0100772110553207 * V,
      0.3223378901422757, 0.0553571815436544 * V, 0.0636610018750175, 0.2696723314583159,
      0.0482142857142857 * V]),
]


def main():
    for name, dim, deg, orbits, p0 in RULES:
        X, w = solve(dim, orbits, deg, np.array(p0, dtype=float))
        print(f"// {len(w)} point{'s' if len(w) > 1 else ''}, exact to degree {deg}")
        print(f"static const double {name}[{len(w)}][4] = {{")
        for x, wt in zip(X, w):
            c = list(x) + [0.0] * (3 - dim)
            print("  {" + ", ".join(f"{v:.17g}" for v in c) + f", {wt:.17g}}},")
        print("};")


if __name__ == "__main__":
    main()
