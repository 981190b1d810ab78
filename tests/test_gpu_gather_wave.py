"""The one-wave persistent gather (one wavefront per chunk, the reference
tensors of the groups with most contributions staged in LDS, the next
chunk's record in flight) against the 4-slot block gather: the same
operations in the same order, so the values must agree bit for bit.  Cases
cover complex and real systems, 16- and 32-bit contribution words (P4
tetrahedra), FP64 and mixed precision, boundary groups read from memory
beside a staged volume group, tensors above the LDS budget (read from
memory), and per-point coefficients (element-matrix planes added first)."""
import numpy as np
import pytest

from _helpers import assert_close

pytestmark = pytest.mark.gpu


def _expr_spec():
    from gfem import workloads as W, x, y
    spec = W.poisson2d(n=9, order=2)
    spec.scalar = "c128"
    omega = spec.field_domain
    spec.terms = [
        dict(kind="bilinear", domain=omega, gauss=4, form=1, coef=1.0 + x() * y()),
        dict(kind="bilinear", domain=omega, gauss=4, form=0, coef=2.0 - 1.0j),
        dict(kind="linear", domain=omega, gauss=4, source=1.0),
    ]
    return spec


def _cases():
    from gfem import workloads as W
    return {
        "helmholtz3d_p2": lambda: W.helmholtz3d(n=6, order=2),
        "helmholtz3d_p1": lambda: W.helmholtz3d(n=9, order=1),
        "helmholtz3d_p3": lambda: W.helmholtz3d(n=3, order=3),
        "poisson2d_p2": lambda: W.poisson2d(n=20, order=2),
        "highorder3d_p4": lambda: W.highorder3d(n=2, order=4, gauss=8),
        "poisson2d_expr": _expr_spec,
    }


def _assemble(monkeypatch, spec, one_wave, precision):
    from gfem import workloads as W
    monkeypatch.setenv("GFEM_GATHER_SLOTS", "4")
    monkeypatch.setenv("GFEM_GATHER_WAVE", "1" if one_wave else "0")
    F = W.build(spec)
    F.pre()
    F.assemble(precision=precision)
    return F


def _bits(a):
    return np.ascontiguousarray(np.asarray(a)).view(np.uint8)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("case", list(_cases()))
def test_one_wave_gather_equals_block_gather(monkeypatch, case, precision):
    print(f"case one-wave gather {case} {precision}", flush=True)
    spec = _cases()[case]()
    Fb = _assemble(monkeypatch, spec, False, precision)
    Fw = _assemble(monkeypatch, spec, True, precision)
    assert not Fb.tuning(precision)["gather_one_wave"]
    assert Fw.tuning(precision)["gather_one_wave"]
    assert np.array_equal(_bits(Fw.values()), _bits(Fb.values())), "A differs from the 4-slot block gather"
    assert np.array_equal(_bits(Fw.rhs(0)), _bits(Fb.rhs(0)))
    # and again: the persistent kernel's state starts clean every launch
    Fw.assemble(precision=precision)
    assert np.array_equal(_bits(Fw.values()), _bits(Fb.values()))


def test_one_wave_gather_matches_oracle(monkeypatch):
    from gfem import workloads as W
    from oracle import oracle as O
    print("case one-wave gather vs oracle helmholtz3d_p2", flush=True)
    spec = W.helmholtz3d(n=4, order=2)
    F = _assemble(monkeypatch, spec, True, "fp64")
    P = O.from_spec(spec)
    P.pre()
    P.assemble()
    assert np.array_equal(F.ai(), P.csr()[0]) and np.array_equal(F.aj(), P.csr()[1])
    assert_close(F.values(), P.values(), 1e-12, "A")


def test_one_wave_gather_after_term_update(monkeypatch):
    """A coefficient update that turns the volume tensor complex changes its
    LDS rows (real 64-byte rows -> complex 128-byte rows): the launch sizes
    the LDS for the new layout and the values still match the block gather."""
    from gfem import workloads as W, ScalarFunction
    print("case one-wave gather after a complex mass coefficient", flush=True)
    spec = W.helmholtz3d(n=5, order=2)
    Fb = _assemble(monkeypatch, spec, False, "fp64")
    Fw = _assemble(monkeypatch, spec, True, "fp64")
    for F in (Fb, Fw):
        F.set_term_coefficient(W.HELMHOLTZ_MASS, ScalarFunction(-30.0 + 4.0j))
        F.assemble()
    assert np.array_equal(_bits(Fw.values()), _bits(Fb.values()))
