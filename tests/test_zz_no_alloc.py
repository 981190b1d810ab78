"""gfem_assemble_into_noalloc (include/gfem.h) through the C ABI: the same
values and right-hand sides as gfem_assemble_into, bit for bit, on two calls in
a row (the second on a smaller system, so the scratch array holds the first
call's data), in FP64 and in mixed precision, after a coefficient change, and
the size and strategy limits it reports."""
import ctypes

import numpy as np
import pytest

import gfem
from gfem import workloads as W

pytestmark = pytest.mark.gpu

GFEM_OK, GFEM_ERR_UNSUPPORTED, GFEM_ERR_TOO_LARGE = 0, 5, 8
FP64, MIXED = 0, 1


@pytest.fixture(scope="module")
def hip():
    lib = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so")
    lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    lib.hipFree.argtypes = [ctypes.c_void_p]
    lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    lib.hipDeviceSynchronize.argtypes = []
    return lib


@pytest.fixture(scope="module")
def abi():
    lib = gfem.load_c_abi()
    for name in ("gfem_assemble_into", "gfem_assemble_into_noalloc"):
        getattr(lib, name).argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                       ctypes.c_void_p]
    lib.gfem_system_info.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int64),
                                     ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int)]
    lib.gfem_status_string.restype = ctypes.c_char_p
    return lib


class DevBuf:
    def __init__(self, hip, nbytes):
        self.hip, self.n = hip, max(int(nbytes), 8)
        self.p = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.p), self.n) == 0

    def host(self):
        out = np.empty(self.n, np.uint8)
        assert self.hip.hipMemcpy(out.ctypes.data, self.p, self.n, 2) == 0  # device to host
        return out

    def __del__(self):
        self.hip.hipFree(self.p)


def _system(spec):
    F = W.build(spec)
    F.pre()
    F.assemble(download=False)  # creates and finalizes the system
    return F


def _both(abi, hip, F, prec):
    """(values, rhs) bytes from gfem_assemble_into and from the no-allocation
    entry point on the same system"""
    h = ctypes.c_void_p(F.system_handle())
    nnz, n, vb = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int()
    assert abi.gfem_system_info(h, prec, ctypes.byref(nnz), ctypes.byref(n), ctypes.byref(vb)) == GFEM_OK
    out = []
    for name in ("gfem_assemble_into", "gfem_assemble_into_noalloc"):
        v = DevBuf(hip, nnz.value * vb.value)
        r = DevBuf(hip, n.value * vb.value * F.num_rhs())
        st = getattr(abi, name)(h, prec, v.p, r.p, None)
        assert st == GFEM_OK, (name, abi.gfem_status_string(st))
        assert hip.hipDeviceSynchronize() == 0
        out.append((v.host(), r.host()))
    return out


@pytest.mark.parametrize("prec", [FP64, MIXED])
def test_noalloc_matches_assemble_into_on_two_calls(abi, hip, prec):
    big = _system(W.helmholtz3d(n=8, order=2))
    small = _system(W.helmholtz3d(n=3, order=1))
    for F in (big, small):
        print(f"case noalloc {F.n_unknown} unknowns prec={prec}", flush=True)
        (v0, r0), (v1, r1) = _both(abi, hip, F, prec)
        assert np.array_equal(v0, v1) and np.array_equal(r0, r1)


def test_noalloc_after_coefficient_change_and_on_a_real_system(abi, hip):
    F = _system(W.helmholtz3d(n=5, order=2))
    for tag, f in W.helmholtz_params(21.0, (0.4, 0.55, 0.6)):
        F.set_term_coefficient(tag, f)
    print("case noalloc after term update", flush=True)
    (v0, r0), (v1, r1) = _both(abi, hip, F, FP64)
    assert np.array_equal(v0, v1) and np.array_equal(r0, r1)
    P = _system(W.poisson2d(n=16, order=3))
    print("case noalloc real poisson P3", flush=True)
    (v0, r0), (v1, r1) = _both(abi, hip, P, FP64)
    assert np.array_equal(v0, v1) and np.array_equal(r0, r1)


def test_noalloc_refuses_what_it_cannot_hold(abi, hip):
    """a system whose scratch exceeds the fixed array, and one with per-point
    element matrices, are refused with status codes; nothing is written"""
    print("case noalloc limits", flush=True)
    v = DevBuf(hip, 8)
    big = _system(W.helmholtz3d(n=60, order=2, jitter=0.0))  # 1.3M tets: ~0.3 GB of factors and loads
    st = abi.gfem_assemble_into_noalloc(ctypes.c_void_p(big.system_handle()), FP64, v.p, v.p, None)
    assert st == GFEM_ERR_TOO_LARGE
    del big
    spec = W.helmholtz3d(n=3, order=1)
    spec.terms[1]["coef"] = -gfem.exp(gfem.x())  # per-point mass coefficient
    F = _system(spec)
    st = abi.gfem_assemble_into_noalloc(ctypes.c_void_p(F.system_handle()), FP64, v.p, v.p, None)
    assert st == GFEM_ERR_UNSUPPORTED
