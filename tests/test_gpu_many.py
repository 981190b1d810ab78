"""gfem_assemble_many: parameter sets of several resident systems assembled
together, one launch per kernel kind per wave, against the same parameters
assembled one system at a time (gfem_term_update + assembly): bit for bit."""
import numpy as np
import pytest

import gfem
from gfem import workloads as W

pytestmark = pytest.mark.gpu


def _systems():
    specs = [W.helmholtz3d(n=3, order=1), W.helmholtz3d(n=3, order=2), W.helmholtz3d(n=2, order=3),
             W.helmholtz3d(n=5, order=1), W.poisson2d(n=12, order=2)]
    Fs = []
    for spec in specs:
        F = W.build(spec)
        F.pre()
        # tile choices tuned in both precisions before either path runs
        F.assemble(precision="mixed", download=False)
        F.assemble(precision="fp64", download=False)
        Fs.append(F)
    return specs, Fs


def _params(rng, spec):
    if spec.name.startswith("poisson"):
        return []
    k = float(rng.uniform(2 * np.pi, 16 * np.pi))
    c = tuple(float(v) for v in rng.uniform(0.25, 0.75, 3))
    return W.helmholtz_params(k, c)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_many_instances_equal_standalone_assembly(precision):
    specs, Fs = _systems()
    rng = np.random.default_rng(11)
    # 14 instances; systems repeat, so the call runs several waves
    order = [0, 1, 2, 3, 4, 0, 0, 2, 1, 3, 0, 2, 4, 1]
    params = [_params(rng, specs[s]) for s in order]
    inst = [Fs[s].many_instance(p) for s, p in zip(order, params)]
    print(f"case assemble_many {len(inst)} instances on {len(Fs)} systems, {precision}", flush=True)
    gfem.assemble_many(inst, precision)
    for j, (s, p) in enumerate(zip(order, params)):
        F = Fs[s]
        for tag, f in p:
            F.set_term_coefficient(tag, f)
        F.assemble(precision=precision)
        ref_a = np.asarray(F.values(), dtype=np.complex128)
        ref_b = np.asarray(F.rhs(0), dtype=np.complex128)
        assert np.array_equal(inst[j].values(precision), ref_a), f"instance {j} (system {s}) A"
        assert np.array_equal(inst[j].rhs(precision), ref_b), f"instance {j} (system {s}) b"


def test_many_leaves_each_system_with_its_last_parameters():
    """After the call each system holds its last instance's parameters: a
    standalone assembly next gives that instance's values."""
    specs, Fs = _systems()
    rng = np.random.default_rng(5)
    order = [0, 0, 1, 0]
    params = [_params(rng, specs[s]) for s in order]
    inst = [Fs[s].many_instance(p) for s, p in zip(order, params)]
    print("case assemble_many last parameters", flush=True)
    gfem.assemble_many(inst, "fp64")
    Fs[0].assemble()
    assert np.array_equal(np.asarray(Fs[0].values(), dtype=np.complex128), inst[3].values())
    Fs[1].assemble()
    assert np.array_equal(np.asarray(Fs[1].values(), dtype=np.complex128), inst[2].values())


def test_many_refuses_per_point_element_matrices():
    spec = W.helmholtz3d(n=3, order=1)
    spec.terms[1]["coef"] = -gfem.exp(gfem.x())  # per-point mass coefficient
    F = W.build(spec)
    F.pre()
    F.assemble(download=False)
    print("case assemble_many refusal", flush=True)
    with pytest.raises(RuntimeError, match="unsupported"):
        gfem.assemble_many([F.many_instance([])], "fp64")
