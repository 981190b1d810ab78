"""Parity at the benchmark sizes, entry by entry against the sequential oracle.

  - 2-D Poisson, 64x64 cells split into 8192 P1 triangles (4225 nodes, interior
    jittered by +-0.2h), Gauss2, Dirichlet u = 0 on the boundary: the case the
    reference's CPU path runs as-is;
  - 3-D Helmholtz, 50^3 cubes x 6 = 750,000 P2 tetrahedra + 30,000 P2 Robin
    triangles (1,030,301 DOFs, 29,096,201 nonzeros), k = 10 pi, Gaussian
    source, complex128: the assembly the benchmark times.

CSR pattern bit-exact (device pre()), FP64 matrix and right-hand side within
1e-12 relative, mixed precision within 2e-5 relative (tests/_helpers)."""
import numpy as np
import pytest

from _helpers import assert_close

pytestmark = pytest.mark.gpu

FP64_RTOL = 1e-12
MIXED_RTOL = 2e-5


def _check(spec):
    from gfem import workloads as W
    from oracle import oracle as O
    F = W.build(spec)
    F.pre()
    P = O.from_spec(spec)
    P.pre()
    P.set_threads(1)  # the sequential checker
    P.assemble()
    ai, aj = P.csr()
    assert (F.n_unknown, F.nnz) == (P.n_unknown, len(aj))
    assert np.array_equal(F.ai(), ai)
    assert np.array_equal(F.aj(), aj)
    F.assemble(precision="fp64")
    assert_close(F.values(), P.values(), FP64_RTOL, spec.name + " A")
    assert_close(F.rhs(0), P.rhs(0), FP64_RTOL, spec.name + " b")
    F.assemble(precision="mixed")
    assert_close(F.values(), P.values(), MIXED_RTOL, spec.name + " A (mixed)")
    assert_close(F.rhs(0), P.rhs(0), MIXED_RTOL, spec.name + " b (mixed)")


def test_poisson2d_64x64_p1():
    from gfem import workloads as W
    print("case full size poisson2d 64x64 P1", flush=True)
    spec = W.poisson2d(n=64, order=1)
    assert sum(b["conn"].shape[0] for b in spec.model.blocks() if b["dim"] == 2) == 8192
    assert spec.model.num_nodes == 4225
    _check(spec)


def test_helmholtz3d_50_p2():
    from gfem import workloads as W
    print("case full size helmholtz3d 50^3x6 P2", flush=True)
    spec = W.helmholtz3d(n=50, order=2)
    _check(spec)
