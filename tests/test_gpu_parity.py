"""HIP engine vs the CPU oracle (same seeded inputs, same CSR pattern).

Bars: DOF numbering, CSR pattern and element maps bit-exact; FP64 matrix and
RHS entries within 1e-12 relative (tests/_helpers.assert_close); mixed
precision (FP32 contraction) within 2e-5 relative; the fused path is
deterministic (bitwise equal on repeated runs)."""
import math

import numpy as np
import pytest

from _helpers import assert_close

pytestmark = pytest.mark.gpu

FP64_RTOL = 1e-12
MIXED_RTOL = 2e-5


def _pair(spec, scalar=None):
    from gfem import workloads as W
    from oracle import oracle as O
    F = W.build(spec, scalar)
    F.pre()
    P = O.from_spec(spec)
    P.pre()
    P.assemble()
    ai, aj = P.csr()
    assert np.array_equal(F.ai(), ai) and np.array_equal(F.aj(), aj)
    return F, P


def _cases():
    from gfem import workloads as W
    return {
        "poisson2d_p1": lambda: W.poisson2d(n=12, order=1),
        "poisson2d_p2": lambda: W.poisson2d(n=7, order=2),
        "poisson2d_p3": lambda: W.poisson2d(n=5, order=3),
        "helmholtz3d_p1": lambda: W.helmholtz3d(n=5, order=1),
        "helmholtz3d_p2": lambda: W.helmholtz3d(n=4, order=2),
        "helmholtz3d_p3": lambda: W.helmholtz3d(n=2, order=3),
        "highorder3d_p4": lambda: W.highorder3d(n=2, order=4, gauss=8),
    }


CASES = ["poisson2d_p1", "poisson2d_p2", "poisson2d_p3", "helmholtz3d_p1", "helmholtz3d_p2", "helmholtz3d_p3",
         "highorder3d_p4"]


@pytest.mark.parametrize("case", CASES)
def test_fused_fp64_matches_oracle(case):
    print(f"case {case} fused fp64", flush=True)
    F, P = _pair(_cases()[case]())
    F.assemble(precision="fp64", strategy="fused")
    assert_close(F.values(), P.values(), FP64_RTOL, case + " A")
    assert_close(F.rhs(0), P.rhs(0), FP64_RTOL, case + " b")


@pytest.mark.parametrize("case", CASES)
def test_reference_strategy_matches_oracle(case):
    print(f"case {case} reference fp64", flush=True)
    F, P = _pair(_cases()[case]())
    F.assemble(precision="fp64", strategy="reference")
    assert_close(F.values(), P.values(), FP64_RTOL, case + " A")
    assert_close(F.rhs(0), P.rhs(0), FP64_RTOL, case + " b")


@pytest.mark.parametrize("case", ["poisson2d_p2", "helmholtz3d_p2", "highorder3d_p4"])
def test_mixed_matches_oracle(case):
    print(f"case {case} fused mixed", flush=True)
    F, P = _pair(_cases()[case]())
    F.assemble(precision="mixed", strategy="fused")
    assert_close(F.values(), P.values(), MIXED_RTOL, case + " A (mixed)")
    assert_close(F.rhs(0), P.rhs(0), MIXED_RTOL, case + " b (mixed)")


def test_fused_is_deterministic():
    from gfem import workloads as W
    print("case determinism helmholtz3d n=6 p2", flush=True)
    F = W.build(W.helmholtz3d(n=6, order=2))
    F.pre()
    F.assemble()
    a1, b1 = F.values().copy(), F.rhs(0).copy()
    for _ in range(3):
        F.assemble()
        assert np.array_equal(F.values(), a1) and np.array_equal(F.rhs(0), b1)


def test_dirichlet_lifting_nonzero_value():
    """u = 1 + 2i on the boundary: b carries A_{i,D} u_D (VectorFactory::addValuesDC)."""
    from gfem import workloads as W
    print("case dirichlet lifting", flush=True)
    spec = W.poisson2d(n=9, order=2)
    spec.dirichlet = [(spec.dirichlet[0][0], 1.0 + 2.0j)]
    spec.scalar = "c128"
    F, P = _pair(spec)
    F.assemble()
    assert np.abs(P.rhs(0)).max() > 0
    assert_close(F.values(), P.values(), FP64_RTOL, "A")
    assert_close(F.rhs(0), P.rhs(0), FP64_RTOL, "b")


def test_multiple_rhs_scalar_product_expression_coefficient():
    from gfem import workloads as W, x, y, sin
    print("case multi-rhs / scalar product / per-point coefficient", flush=True)
    spec = W.poisson2d(n=8, order=2)
    spec.scalar = "c128"
    omega = spec.field_domain
    spec.terms = [
        dict(kind="bilinear", domain=omega, gauss=4, form=1, coef=1.0 + x() * y()),
        dict(kind="bilinear", domain=omega, gauss=4, form=0, coef=2.0 - 1.0j, tf_coef=0.5 + 0.25j, hermitian=False),
        dict(kind="linear", domain=omega, gauss=4, source=sin(x() * 3.0) + 1j, rhs_idx=0),
        dict(kind="linear", domain=omega, gauss=3, source=-2.0, tf_coef=1j, rhs_idx=1),
    ]
    F, P = _pair(spec)
    for strat in ("fused", "reference", "dense"):
        F.assemble(strategy=strat)
        assert_close(F.values(), P.values(), FP64_RTOL, strat + " A")
        for r in range(2):
            assert_close(F.rhs(r), P.rhs(r), FP64_RTOL, f"{strat} b{r}")


def test_frequency_module_dt_terms():
    from gfem import workloads as W
    print("case AFrequency module", flush=True)
    spec = W.helmholtz3d(n=3, order=2)
    spec.frequency = 3.0
    spec.terms.append(dict(kind="bilinear", domain=[(3, 1)], gauss=4, form=0, coef=0.7, dt=2))
    spec.terms.append(dict(kind="bilinear", domain=[(2, 1), (2, 2)], gauss=4, form=0, coef=0.3, dt=1))
    F, P = _pair(spec)
    F.assemble()
    assert_close(F.values(), P.values(), FP64_RTOL, "A")
    assert_close(F.rhs(0), P.rhs(0), FP64_RTOL, "b")


def test_symmetric_storage():
    from gfem import workloads as W
    print("case symmetric storage", flush=True)
    spec = W.poisson2d(n=10, order=2)
    spec.symmetric = True
    F, P = _pair(spec)
    F.assemble()
    assert_close(F.values(), P.values(), FP64_RTOL, "A")
    assert_close(F.rhs(0), P.rhs(0), FP64_RTOL, "b")


def test_compiled_coefficients_match_interpreter(monkeypatch):
    """Run-time compiled coefficient programs (hipRTC) and the interpreter
    give the same load vector; both match the oracle."""
    from gfem import workloads as W
    print("case compiled vs interpreted coefficient programs", flush=True)
    spec = W.helmholtz3d(n=4, order=2)
    F1, P = _pair(spec)
    F1.assemble()
    monkeypatch.setenv("GFEM_JIT", "0")
    F2 = W.build(spec)
    F2.pre()
    F2.assemble()
    assert_close(F1.rhs(0), P.rhs(0), FP64_RTOL, "compiled b")
    assert_close(F2.rhs(0), P.rhs(0), FP64_RTOL, "interpreted b")
    assert_close(F1.rhs(0), F2.rhs(0), 1e-14, "compiled vs interpreted")



@pytest.mark.parametrize("slots", [2, 8])
@pytest.mark.parametrize("case", ["helmholtz3d_p2", "poisson2d_p3", "highorder3d_p4"])
@pytest.mark.parametrize("precision,rtol", [("fp64", FP64_RTOL), ("mixed", MIXED_RTOL)])
def test_gather_slots_per_thread(monkeypatch, case, precision, rtol, slots):
    """The gather with 2 or 8 slots per thread (256- or 64-thread blocks)
    assembles the same matrix as the default 4 (128-thread blocks), to the
    same bars (16-bit words for P2/P3 triangles, 32-bit words for P4
    tetrahedra)."""
    from gfem import workloads as W
    print(f"case gather {slots} slots {case} {precision}", flush=True)
    spec = _cases()[case]()
    monkeypatch.setenv("GFEM_GATHER_SLOTS", "4")
    F1, P = _pair(spec)
    F1.assemble(precision=precision)
    monkeypatch.setenv("GFEM_GATHER_SLOTS", str(slots))
    F2 = W.build(spec)
    F2.pre()
    F2.assemble(precision=precision)
    assert F2.tuning(precision)["gather_slots"] == slots
    assert_close(F2.values(), P.values(), rtol, f"{slots}-slot A")
    assert_close(F2.rhs(0), P.rhs(0), rtol, f"{slots}-slot b")
    assert_close(F2.values(), F1.values(), 1e-14 if precision == "fp64" else rtol, f"{slots} vs 4 slots")


def test_systems_on_two_meshes_assemble_in_any_order():
    """Two formulations on different meshes, uploaded one after the other,
    then assembled alternately: each keeps its own nodes."""
    from gfem import workloads as W
    print("case two meshes interleaved", flush=True)
    s1, s2 = W.helmholtz3d(n=5, order=1), W.poisson2d(n=9, order=2)
    F1, P1 = _pair(s1)
    F2, P2 = _pair(s2)
    F1.assemble()
    F2.assemble()
    F1.assemble()
    assert_close(F1.values(), P1.values(), FP64_RTOL, "mesh 1 A")
    assert_close(F1.rhs(0), P1.rhs(0), FP64_RTOL, "mesh 1 b")
    F2.assemble()
    assert_close(F2.values(), P2.values(), FP64_RTOL, "mesh 2 A")


DENSE_FP32_RTOL = 1e-4  # FP32 products summed over 7Q terms in the GEMM


@pytest.mark.parametrize("case", ["poisson2d_p2", "helmholtz3d_p2", "highorder3d_p4", "poisson2d_p3"])
def test_dense_strategy_matches_oracle(case):
    """Every bilinear term as per-point element matrices (one GEMM per group
    on the matrix cores): FP64 to the FP64 bar, FP32 GEMM to its own bar."""
    print(f"case dense strategy {case}", flush=True)
    F, P = _pair(_cases()[case]())
    F.assemble(precision="fp64", strategy="dense")
    assert_close(F.values(), P.values(), FP64_RTOL, case + " A (dense fp64)")
    assert_close(F.rhs(0), P.rhs(0), FP64_RTOL, case + " b (dense fp64)")
    F.assemble(precision="mixed", strategy="dense")
    assert_close(F.values(), P.values(), DENSE_FP32_RTOL, case + " A (dense fp32)")
    # the fused strategy afterwards is unaffected by the dense buffers
    F.assemble(precision="fp64", strategy="fused")
    assert_close(F.values(), P.values(), FP64_RTOL, case + " A (fused after dense)")


# three bf16 products per FP32 one: each operand is hi + lo to ~2^-18, lo*lo is
# dropped (~2^-18), so each product is good to ~3 * 2^-18 ~ 1e-5; Helmholtz
# entries cancel the stiffness against k^2 times the mass, which amplifies that
# ~30x on the worst entries (measured 4.4e-4 on helmholtz3d_p2)
BF16X3_RTOL = 1e-3


@pytest.mark.parametrize("case", ["helmholtz3d_p2", "highorder3d_p4", "poisson2d_p3"])
def test_bf16x3_gemm_matches_oracle(case):
    """The element GEMM on the bf16 matrix cores with split operands: within
    its bar of the oracle (bf16 alone, lo terms dropped, is ~2^-8 per product
    and fails it), and within the same bar of the FP32 GEMM; FP32 again after
    switching back gives the FP32 values bit for bit."""
    print(f"case bf16x3 gemm {case}", flush=True)
    F, P = _pair(_cases()[case]())
    F.assemble(precision="mixed", strategy="dense")
    v32 = F.values()
    F.set_gemm_mode("bf16x3")
    F.assemble(precision="mixed", strategy="dense")
    vb = F.values()
    assert_close(vb, P.values(), BF16X3_RTOL, case + " A (dense bf16x3)")
    assert_close(F.rhs(0), P.rhs(0), DENSE_FP32_RTOL, case + " b (dense bf16x3)")
    assert_close(vb, v32, BF16X3_RTOL, case + " bf16x3 vs fp32 GEMM")
    assert not np.array_equal(vb, v32)  # the split path ran
    F.set_gemm_mode("fp32")
    F.assemble(precision="mixed", strategy="dense")
    assert np.array_equal(F.values(), v32)


def test_per_point_gemm_matches_direct_kernel(monkeypatch):
    """Per-point bilinear coefficients on P3 triangles: the GEMM form and the
    direct per-point kernel (GFEM_DENSE_GEMM=0) agree with the oracle."""
    from gfem import workloads as W, x, y
    print("case per-point gemm vs direct", flush=True)
    spec = W.poisson2d(n=6, order=3)
    spec.terms[0] = dict(kind="bilinear", domain=spec.field_domain, gauss=6, form=1, coef=2.0 + x() * x() - y())
    spec.terms.append(dict(kind="bilinear", domain=spec.field_domain, gauss=6, form=0, coef=1.0 + y()))
    F1, P = _pair(spec)
    F1.assemble()
    monkeypatch.setenv("GFEM_DENSE_GEMM", "0")
    F2 = W.build(spec)
    F2.pre()
    F2.assemble()
    assert_close(F1.values(), P.values(), FP64_RTOL, "gemm A")
    assert_close(F2.values(), P.values(), FP64_RTOL, "direct A")
    F1.assemble(precision="mixed")
    assert_close(F1.values(), P.values(), DENSE_FP32_RTOL, "gemm A (fp32)")
    # every term as element matrices, with the A rows read from memory
    F1.assemble(precision="fp64", strategy="dense")
    assert_close(F1.values(), P.values(), FP64_RTOL, "dense gemm A (fp64, read rows)")
    F1.set_gemm_mode("bf16x3")
    F1.assemble(precision="mixed", strategy="dense")
    assert_close(F1.values(), P.values(), BF16X3_RTOL, "dense gemm A (bf16x3, read rows)")


@pytest.mark.parametrize("gemm", ["fp64", "fp32", "bf16x3"])
def test_generated_and_read_gemm_rows_agree(gemm):
    """The element-matrix kernel generates A from the element factors when
    every coefficient is constant and reads k_coef_rows' rows otherwise.  The
    same complex Helmholtz problem with its mass coefficient written as an
    expression (-k^2 + 0 x) takes the read path: both agree with the oracle
    to the arithmetic's bar, and with each other."""
    from gfem import workloads as W, x
    print(f"case generated vs read gemm rows {gemm}", flush=True)
    spec = W.helmholtz3d(n=4, order=2)
    k2 = spec.terms[1]["coef"]
    F1, P = _pair(spec)
    spec_x = W.helmholtz3d(n=4, order=2)
    spec_x.terms[1]["coef"] = k2 + 0.0 * x()
    F2 = W.build(spec_x)
    F2.pre()
    prec = "fp64" if gemm == "fp64" else "mixed"
    tol = {"fp64": FP64_RTOL, "fp32": DENSE_FP32_RTOL, "bf16x3": BF16X3_RTOL}[gemm]
    vals = []
    for F in (F1, F2):
        F.assemble(download=False)  # creates the system
        F.set_gemm_mode("bf16x3" if gemm == "bf16x3" else "fp32")
        F.assemble(precision=prec, strategy="dense")
        assert_close(F.values(), P.values(), tol, f"{gemm} A")
        vals.append(F.values())
    assert_close(vals[1], vals[0], tol, f"{gemm} read vs generated")


def test_autotuned_assembly_matches_untuned(monkeypatch):
    """The autotuner's choices are in their candidate sets and the tuned
    assembly equals the untuned one (GFEM_AUTOTUNE=0) to round-off."""
    from gfem import workloads as W
    print("case autotune", flush=True)
    spec = W.helmholtz3d(n=6, order=2)
    F1, P = _pair(spec)
    F1.assemble(precision="mixed")
    F1.assemble()  # FP64 last: values() below are the FP64 assembly
    for prec in ("fp64", "mixed"):
        t = F1.tuning(prec)
        assert t["gather_slots"] in (2, 4, 8) and t["linear_eb_shift"] in (1, 2, 3), t
        assert t["rhs_lanes"] in (2, 4, 8), t
    monkeypatch.setenv("GFEM_AUTOTUNE", "0")
    F2 = W.build(spec)
    F2.pre()
    F2.assemble()
    assert F2.tuning("fp64") == dict(gather_slots=4, linear_eb_shift=2, spmv_lanes=8, rhs_lanes=2,
                                     gather_one_wave=False)
    assert_close(F1.values(), F2.values(), 1e-14, "tuned vs untuned A")
    assert_close(F1.rhs(0), F2.rhs(0), 1e-14, "tuned vs untuned b")
    assert_close(F1.values(), P.values(), FP64_RTOL, "tuned A")


def test_element_load_layout_follows_term_updates():
    """Element loads are stored real while every linear term is real, complex
    otherwise.  Two right-hand sides, each with a source on one group only
    (triangles for b0, boundary curves for b1): a coefficient update that
    makes the b0 source complex, then real again, gives the oracle's b0 and
    b1 each time (a group's missing right-hand side reads zeros in either
    layout)."""
    from gfem import ScalarFunction, workloads as W
    from oracle import oracle as O
    print("case element-load layout flips", flush=True)
    spec = W.poisson2d(n=6, order=2)
    spec.scalar = "c128"
    spec.dirichlet = []
    omega, gamma = spec.field_domain, [(1, 1), (1, 2), (1, 3), (1, 4)]
    spec.terms = [dict(kind="bilinear", domain=omega, gauss=4, form=1),
                  dict(kind="linear", domain=omega, gauss=4, source=1.0, rhs_idx=0),
                  dict(kind="linear", domain=gamma, gauss=4, source=2.0, rhs_idx=1)]
    F = W.build(spec)
    F.pre()
    for c in (1.0, 1.0 + 2.0j, 3.0, -0.5j):
        spec.terms[1]["source"] = c
        F.set_term_coefficient(1, ScalarFunction(c))
        P = O.from_spec(spec)
        P.pre()
        P.assemble()
        assert np.abs(P.rhs(1)).max() > 0 and np.abs(P.rhs(0)).max() > 0
        for strat in ("fused", "reference"):
            F.assemble(strategy=strat)
            assert_close(F.values(), P.values(), FP64_RTOL, f"A c={c} {strat}")
            for r in range(2):
                assert_close(F.rhs(r), P.rhs(r), FP64_RTOL, f"b{r} c={c} {strat}")


def _degenerate_specs():
    from gfem import workloads as W
    allfixed = W.poisson2d(n=1, order=1)            # 2 triangles, every node on the Dirichlet boundary
    linonly = W.poisson2d(n=4, order=2)             # a source and no bilinear term: no nonzeros
    linonly.dirichlet = []
    linonly.terms = [linonly.terms[1]]
    bilonly = W.poisson2d(n=4, order=2)             # no source: b is the Dirichlet lift alone
    bilonly.dirichlet = [(bilonly.dirichlet[0][0], 2.0)]
    bilonly.terms = [bilonly.terms[0]]
    return {"all_fixed": allfixed, "linear_only": linonly, "bilinear_only": bilonly}


@pytest.mark.parametrize("device", [False, True])
@pytest.mark.parametrize("case", ["all_fixed", "linear_only", "bilinear_only"])
def test_degenerate_systems_match_oracle(case, device):
    """Edge cases of the assembly: no unknown DOF at all (empty A and b), a
    formulation without bilinear terms (empty pattern, b only), and one
    without linear terms (b = the Dirichlet lift).  Every strategy and the
    mixed path give the oracle's (empty) arrays."""
    from oracle import oracle as O
    from gfem import workloads as W
    print(f"case degenerate {case} device pre={device}", flush=True)
    spec = _degenerate_specs()[case]
    P = O.from_spec(spec)
    P.pre()
    P.assemble()
    F = W.build(spec)
    F.pre(device=device)
    assert (F.n_unknown, F.nnz) == (P.n_unknown, len(P.values()))
    for strat in ("fused", "reference", "dense"):
        F.assemble(strategy=strat)
        assert F.values().shape == P.values().shape and F.rhs(0).shape == P.rhs(0).shape
        assert_close(F.values(), P.values(), FP64_RTOL, f"{case} A {strat}")
        assert_close(F.rhs(0), P.rhs(0), FP64_RTOL, f"{case} b {strat}")
    F.assemble(precision="mixed")
    assert_close(F.values(), P.values(), MIXED_RTOL, f"{case} A mixed")
    assert_close(F.rhs(0), P.rhs(0), MIXED_RTOL, f"{case} b mixed")
