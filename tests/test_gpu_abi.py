"""The C ABI (include/gfem.h) called directly through ctypes on the GPU: the
argument checks a caller in Formulation::assemble relies on."""
import ctypes

import numpy as np
import pytest

import gfem

pytestmark = pytest.mark.gpu

_i64p = ctypes.POINTER(ctypes.c_int64)
_i32p = ctypes.POINTER(ctypes.c_int32)
_dp = ctypes.POINTER(ctypes.c_double)

GFEM_OK, GFEM_ERR_INVALID_ARG, GFEM_ERR_UNSUPPORTED = 0, 1, 5
AT_FIXED, AT_UNKNOWN, AT_GLOBAL, AT_LINKED, AT_LINKED_BUBBLE = 1, 2, 4, 5, 6


def _p(a, t):
    return a.ctypes.data_as(t)


@pytest.fixture(scope="module")
def abi():
    lib = gfem.load_c_abi()
    ctx = ctypes.c_void_p()
    assert lib.gfem_ctx_create(0, ctypes.byref(ctx)) == GFEM_OK
    lib.gfem_ctx_last_error.restype = ctypes.c_char_p
    yield lib, ctx
    lib.gfem_ctx_destroy(ctx)


def _two_triangles(lib, ctx):
    """unit square split into two P1 triangles, four unknown DOFs, full pattern"""
    xyz = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], np.float64)
    assert lib.gfem_nodes_upload(ctx, ctypes.c_int64(4), _p(xyz, _dp)) == GFEM_OK
    ai = np.array([0, 4, 8, 12, 16], np.int64)
    aj = np.tile(np.arange(4, dtype=np.int64), 4)
    sysh = ctypes.c_void_p()
    st = lib.gfem_system_create(ctx, 0, ctypes.c_int64(4), _p(ai, _i64p), _p(aj, _i64p), 0, ctypes.c_int64(0), None,
                                1, ctypes.c_double(0.0), ctypes.byref(sysh))
    assert st == GFEM_OK, lib.gfem_ctx_last_error(ctx)
    conn = np.array([[1, 2, 3], [1, 3, 4]], np.int64)
    num = conn - 1
    e2n = np.array([[4 * r + c for r in e for c in e] for e in num], np.int64)
    return sysh, conn, num, e2n


@pytest.mark.parametrize("bad,expect", [(AT_GLOBAL, GFEM_ERR_UNSUPPORTED), (AT_LINKED, GFEM_ERR_UNSUPPORTED),
                                        (AT_LINKED_BUBBLE, GFEM_ERR_UNSUPPORTED), (9, GFEM_ERR_INVALID_ARG)])
def test_group_add_refuses_linked_and_global_keys(abi, bad, expect):
    """Linked and global keys (Dof.h:27-35) are refused with a status code
    instead of having their contributions dropped; the same group with plain
    unknown keys is accepted."""
    lib, ctx = abi
    print(f"case group_add assemble type {bad}", flush=True)
    sysh, conn, num, e2n = _two_triangles(lib, ctx)
    try:
        typ = np.full(conn.shape, AT_UNKNOWN, np.int32)
        typ[1, 2] = bad
        g = ctypes.c_int()
        st = lib.gfem_group_add(sysh, 2, 2, ctypes.c_int64(2), _p(conn, _i64p), _p(num, _i64p), _p(typ, _i32p),
                                _p(e2n, _i64p), ctypes.byref(g))
        assert st == expect, lib.gfem_ctx_last_error(ctx)
        typ[1, 2] = AT_UNKNOWN
        st = lib.gfem_group_add(sysh, 2, 2, ctypes.c_int64(2), _p(conn, _i64p), _p(num, _i64p), _p(typ, _i32p),
                                _p(e2n, _i64p), ctypes.byref(g))
        assert st == GFEM_OK, lib.gfem_ctx_last_error(ctx)
    finally:
        lib.gfem_system_destroy(sysh)
