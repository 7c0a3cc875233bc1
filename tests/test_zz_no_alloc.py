"""pa_pinv_solve_noalloc (the pseudo-inverse solve with its device buffers in a static per-GPU
array, for callers that cannot allocate) against pa_pinv_solve, through the C ABI: bit-identical
x on all three kernel paths, on two calls in a row (the second smaller, so stale contents of the
array from the first would show), and the size limit reported before any device work."""
import ctypes

import numpy as np
import pytest

from plasticity_amd import native

pytestmark = pytest.mark.gpu


def _systems(rng, count, n):
    """Dense n x n systems: a third of them with one exact zero singular value (the pseudo-inverse
    branch), the rest well conditioned."""
    A = np.empty((count, n, n))
    for e in range(count):
        U, _ = np.linalg.qr(rng.standard_normal((n, n)))
        V, _ = np.linalg.qr(rng.standard_normal((n, n)))
        s = rng.uniform(0.5, 2.0, n)
        if e % 3 == 0:
            s[rng.integers(n)] = 0.0
        A[e] = (U * s) @ V.T
    b = rng.standard_normal((count, n))
    return np.ascontiguousarray(A), np.ascontiguousarray(b)


def _call(fn, path, A, b):
    count, n = A.shape[0], A.shape[1]
    x = np.full((count, n), np.nan)
    P = ctypes.c_void_p
    rc = fn(ctypes.c_int(path), ctypes.c_int(n), ctypes.c_longlong(count), P(A.ctypes.data), P(b.ctypes.data), P(x.ctypes.data))
    assert rc == 0, (rc, native.lib().pa_last_error())
    return x


@pytest.mark.parametrize("n,count", [(12, 700), (18, 300), (48, 60)])
@pytest.mark.parametrize("path", [0, 1, 2])
def test_noalloc_equals_pa_pinv_solve_on_two_calls(path, n, count):
    lib = native.lib()
    rng = np.random.default_rng(1000 * path + n)
    for k, cnt in enumerate((count, count // 3 + 1)):  # second call smaller, with other data
        print("no-alloc pinv: path %d, n %d, call %d, %d systems" % (path, n, k, cnt), flush=True)
        A, b = _systems(rng, cnt, n)
        x_ref = _call(lib.pa_pinv_solve, path, A, b)
        x = _call(lib.pa_pinv_solve_noalloc, path, A, b)
        assert np.array_equal(x, x_ref), (path, n, k, np.abs(x - x_ref).max())
        # and it solves the well-conditioned systems (the singular ones follow the reference's
        # rule for one small singular value, which may invert it: no library routine to compare)
        ok = np.arange(cnt) % 3 != 0
        x_np = np.linalg.solve(A[ok], b[ok][:, :, None])[:, :, 0]
        assert np.abs(x[ok] - x_np).max() < 1e-10 * max(1.0, np.abs(x_np).max())


def test_noalloc_rejects_more_systems_than_its_array_holds():
    lib = native.lib()
    n, count = 48, 5000  # 5000 x ~38 KB > 32 MiB
    A = np.zeros((count, n, n))
    b = np.zeros((count, n))
    x = np.zeros((count, n))
    P = ctypes.c_void_p
    rc = lib.pa_pinv_solve_noalloc(ctypes.c_int(0), ctypes.c_int(n), ctypes.c_longlong(count), P(A.ctypes.data),
                                   P(b.ctypes.data), P(x.ctypes.data))
    assert rc == native.PA_ERR_ARG
    msg = lib.pa_last_error().decode()
    assert "exceed the scratch array" in msg and "at most 873 systems" in msg, msg
