"""The batched complex GEMM (every tile shape, both precisions) against
numpy, including the triangular-operand and lower-output variants the
solver uses."""
import numpy as np
import pytest

from tbk import _lib

pytestmark = pytest.mark.gpu

SHAPES = [(288, 288, 288), (37, 300, 65), (96, 96, 32), (130, 70, 17), (9, 9, 9)]


def run(prec, tile, m, n, k, opa, opb, tri, A, B):
    L = _lib.lib()
    dt = np.complex64 if prec else np.complex128
    a = np.asfortranarray(A, dtype=dt)
    b = np.asfortranarray(B, dtype=dt)
    c = np.zeros((m, n), dtype=dt, order="F")
    rc = L.tbk_debug_zgemm(0, prec, tile, m, n, k, opa, opb, tri, a.ctypes.data, a.shape[0],
                           b.ctypes.data, b.shape[0], c.ctypes.data, m)
    assert rc == 0
    return c


def rand(rng, r, c):
    return rng.standard_normal((r, c)) + 1j * rng.standard_normal((r, c))


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("tile", [0, 1, 2, 3])
def test_gemm_tiles_all_ops(prec, tile):
    rng = np.random.default_rng(10 + tile)
    tol = 1e-13 if prec == 0 else 3e-5
    for (m, n, k) in SHAPES:
        for opa in (0, 1):
            for opb in (0, 1):
                A = rand(rng, m, k) if opa == 0 else rand(rng, k, m)
                B = rand(rng, k, n) if opb == 0 else rand(rng, n, k)
                opA = A if opa == 0 else A.conj().T
                opB = B if opb == 0 else B.conj().T
                ref = opA @ opB
                c = run(prec, tile, m, n, k, opa, opb, 0, A, B)
                err = np.abs(c - ref).max() / (np.sqrt(k) * max(1.0, np.abs(ref).max()))
                assert err < tol, (prec, tile, m, n, k, opa, opb, err)


@pytest.mark.parametrize("tile", [0, 1, 2, 3])
def test_gemm_triangular_variants(tile):
    rng = np.random.default_rng(3)
    n = 150
    U = np.triu(rand(rng, n, n))
    Lw = np.tril(rand(rng, n, n))
    X = rand(rng, n, n)
    # tri 1: op(A) upper
    c = run(0, tile, n, n, n, 0, 0, 1, U, X)
    assert np.abs(c - U @ X).max() < 1e-11
    # tri 2: op(A) lower (A = U^H)
    c = run(0, tile, n, n, n, 1, 0, 2, U, X)
    assert np.abs(c - U.conj().T @ X).max() < 1e-11
    # tri 3: op(B) upper
    c = run(0, tile, n, n, n, 0, 0, 3, X, U)
    assert np.abs(c - X @ U).max() < 1e-11
    # tri 4: lower triangle of C only
    c = run(0, tile, n, n, n, 0, 0, 4, X, Lw)
    ref = X @ Lw
    low = np.tril(np.ones((n, n), bool))
    assert np.abs((c - ref)[low]).max() < 1e-11
    # tri 5: lower triangle of Y^H X mirrored into the upper one (exactly
    # Hermitian result, the upper triangle the conjugate of the lower)
    for prec, tol in ((0, 1e-11), (1, 2e-3)):
        c = run(prec, tile, n, n, n, 1, 0, 5, X, Lw)
        ref = X.conj().T @ Lw
        herm = np.tril(ref) + np.tril(ref, -1).conj().T
        assert np.abs(c - herm).max() < tol, (prec, np.abs(c - herm).max())
        off = ~np.eye(n, dtype=bool)
        assert np.array_equal(c[off], c.conj().T[off])
    # tri 7: upper triangle of C only
    c = run(0, tile, n, n, n, 0, 0, 7, X, Lw)
    ref = X @ Lw
    up = np.triu(np.ones((n, n), bool))
    assert np.abs((c - ref)[up]).max() < 1e-11
    # tri 6: op(A) lower (A = U^H) with the Hermitian mirror of tri 5
    c = run(0, tile, n, n, n, 1, 0, 6, U, X)
    ref = U.conj().T @ X
    herm = np.tril(ref) + np.tril(ref, -1).conj().T
    assert np.abs(c - herm).max() < 1e-11
    # tri 9: a full Hermitian A read as L = tril(A, -1) + diag(A) / 2, whatever
    # its upper triangle holds; then A = L + L^H (the pressure's products)
    H = X + X.conj().T
    Hs = np.tril(H) + np.triu(rand(rng, n, n), 1)          # the upper part must not be read
    Lh = np.tril(H, -1) + np.diag(np.diag(H)) / 2
    for m2 in (n, 37):
        c = run(0, tile, n, m2, n, 0, 0, 9, Hs, X[:, :m2])
        assert np.abs(c - Lh @ X[:, :m2]).max() < 1e-11
        quad = 2 * np.einsum("ij,ij->j", X[:, :m2].conj(), c).real
        want = np.einsum("ij,ij->j", X[:, :m2].conj(), H @ X[:, :m2]).real
        assert np.abs(quad - want).max() < 1e-9 * np.abs(want).max()
