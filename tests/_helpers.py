"""Shared comparison helpers for the parity tests."""
import numpy as np


def assert_close(got, ref, rtol, name="", floor=1e-2):
    """Entrywise |got - ref| <= rtol * (|ref| + floor * max|ref|).

    The absolute floor (floor * rtol * max|ref|) covers entries that cancel to
    near zero in the sum over elements, whose relative error is not
    meaningful; everything else is held to rtol relative."""
    got = np.asarray(got, dtype=np.complex128)
    ref = np.asarray(ref, dtype=np.complex128)
    assert got.shape == ref.shape, f"{name}: shape {got.shape} vs {ref.shape}"
    if ref.size == 0:
        return
    scale = float(np.max(np.abs(ref)))
    err = np.abs(got - ref)
    lim = rtol * (np.abs(ref) + floor * scale)
    bad = err > lim
    if bad.any():
        i = int(np.argmax(err - lim))
        raise AssertionError(f"{name}: {int(bad.sum())}/{ref.size} entries off; worst #{i}: {got[i]} vs {ref[i]} "
                             f"(rtol {rtol}, scale {scale:.3e})")


def csr_to_dense(ai, aj, v, n):
    A = np.zeros((n, n), dtype=np.complex128)
    for r in range(n):
        A[r, aj[ai[r]:ai[r + 1]]] += v[ai[r]:ai[r + 1]]
    return A
