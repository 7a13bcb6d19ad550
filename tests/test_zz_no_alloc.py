"""tbk_zhegv_noalloc (workspace in a fixed device array, no allocation)
against tbk_zhegv (the same solve on allocated memory), through the C ABI.

The two run the same kernels on the same inputs, so eigenvalues and
eigenvectors must agree bit for bit; two calls in a row, the second at a
smaller order, check that the array's stale contents from the first call do
not leak into the second.
"""
import ctypes as C

import numpy as np
import pytest

from tbk import _lib


def pencil(n, seed):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    b = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    return a + a.conj().T, b @ b.conj().T / n + np.eye(n)


def call(fn, h, s):
    n = h.shape[0]
    hc = np.ascontiguousarray(h.T, np.complex128)     # column-major
    sc = np.ascontiguousarray(s.T, np.complex128)
    w = np.zeros(n)
    z = np.zeros(n * n, np.complex128)
    rc = fn(0, n, hc.ctypes.data, sc.ctypes.data, w.ctypes.data, z.ctypes.data)
    return rc, w, z.reshape(n, n).T


def test_noalloc_rejects_orders_beyond_its_array():
    L = _lib.lib()
    h = np.zeros(2, np.complex128)
    w = np.zeros(2)
    for n in (1, 513):
        assert L.tbk_zhegv_noalloc(0, n, h.ctypes.data, h.ctypes.data, w.ctypes.data, None) == _lib.TBK_ERR_ARG
    assert L.tbk_zhegv_noalloc(-1, 4, h.ctypes.data, h.ctypes.data, w.ctypes.data, None) == _lib.TBK_ERR_ARG


@pytest.mark.gpu
def test_noalloc_matches_allocating_entry_point():
    L = _lib.lib()
    for n, seed in ((512, 11), (97, 12)):
        print(f"case: zhegv_noalloc n={n} seed={seed}", flush=True)
        h, s = pencil(n, seed)
        rc0, w0, z0 = call(L.tbk_zhegv, h, s)
        rc1, w1, z1 = call(L.tbk_zhegv_noalloc, h, s)
        assert rc0 == 0 and rc1 == 0, (rc0, rc1)
        res = float(np.abs(h @ z1 - s @ z1 * w1).max()) / max(1.0, np.abs(w1).max())
        print(f"n={n}: max|dw|={np.abs(w1 - w0).max():.1e} max|dz|={np.abs(z1 - z0).max():.1e} "
              f"residual={res:.1e}", flush=True)
        assert np.array_equal(w0, w1)
        assert np.array_equal(z0, z1)
        assert res < 1e-11


@pytest.mark.gpu
def test_noalloc_indefinite_overlap_is_an_error():
    L = _lib.lib()
    h = np.eye(4, dtype=complex)
    s = np.diag([1.0, 1.0, -0.5, 1.0]).astype(complex)
    rc, _, _ = call(L.tbk_zhegv_noalloc, h, s)
    assert rc == _lib.TBK_ERR_OVERLAP
    h, s = pencil(40, 13)          # a good pencil after the failure
    rc0, w0, _ = call(L.tbk_zhegv, h, s)
    rc1, w1, _ = call(L.tbk_zhegv_noalloc, h, s)
    assert rc0 == 0 and rc1 == 0 and np.array_equal(w0, w1)
