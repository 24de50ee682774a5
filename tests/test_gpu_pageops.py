"""GPU checks of the page-batched primitives through the C ABI
(include/orf_pageops.h) against numpy/LAPACK in float64."""
import ctypes as C

import numpy as np
import pytest

from opticrcwafdtd_amd import _lib

pytestmark = pytest.mark.gpu


def _crand(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def _f(a):
    """Pages last -> column-major interleaved buffer (batch, n, m) F-ordered per page."""
    return np.ascontiguousarray(np.transpose(a, (0, 2, 1)))


@pytest.mark.parametrize("opA,opB,m,n,k,batch", [
    ("N", "N", 17, 23, 9, 3), ("N", "N", 98, 98, 98, 2), ("T", "N", 64, 130, 70, 2),
    ("N", "C", 242, 61, 100, 2), ("C", "T", 33, 65, 129, 1),
    # few columns: the row-streaming product (4 and 8 column variants, rows past 256),
    # and a long inner dimension that goes back to the MFMA tiles
    ("N", "N", 242, 3, 242, 3), ("N", "N", 300, 7, 61, 2), ("N", "N", 50, 2, 600, 1)])
def test_pagemtimes(opA, opB, m, n, k, batch):
    rng = np.random.default_rng(1)
    A = _crand(rng, batch, m if opA == "N" else k, k if opA == "N" else m)
    B = _crand(rng, batch, k if opB == "N" else n, n if opB == "N" else k)
    op = {"N": lambda x: x, "T": lambda x: np.swapaxes(x, -1, -2), "C": lambda x: np.conj(np.swapaxes(x, -1, -2))}
    ref = op[opA](A) @ op[opB](B)
    Af, Bf = _f(A), _f(B)
    Cf = np.zeros((batch, n, m), complex)
    lib = _lib.load_library()
    rc = lib.orf_pagemtimes(_lib.context(), opA.encode(), opB.encode(), m, n, k, _lib.dptr(Af), A[0].size,
                            _lib.dptr(Bf), B[0].size, _lib.dptr(Cf), batch)
    assert rc == 0, _lib.last_error()
    got = np.transpose(Cf, (0, 2, 1))
    assert np.max(np.abs(got - ref)) < 1e-12 * max(1.0, np.max(np.abs(ref)))


@pytest.mark.parametrize("opA,opB,m,n,k,batch", [
    ("N", "N", 17, 23, 9, 3), ("N", "N", 242, 242, 242, 2), ("T", "N", 64, 130, 70, 2), ("N", "C", 121, 61, 100, 2)])
def test_pagemtimes_f32(opA, opB, m, n, k, batch):
    """FP32 products with FP32 accumulation: exact up to the FP32 accumulation bound of the
    complex64-rounded operands, and visibly different from the FP64 product."""
    rng = np.random.default_rng(3)
    A = _crand(rng, batch, m if opA == "N" else k, k if opA == "N" else m)
    B = _crand(rng, batch, k if opB == "N" else n, n if opB == "N" else k)
    op = {"N": lambda x: x, "T": lambda x: np.swapaxes(x, -1, -2), "C": lambda x: np.conj(np.swapaxes(x, -1, -2))}
    A32, B32 = A.astype(np.complex64).astype(complex), B.astype(np.complex64).astype(complex)
    ref32 = op[opA](A32) @ op[opB](B32)
    bound = op[opA](np.abs(A32)) @ op[opB](np.abs(B32))
    Cf = np.zeros((batch, n, m), complex)
    lib = _lib.load_library()
    rc = lib.orf_pagemtimes_f32(_lib.context(), opA.encode(), opB.encode(), m, n, k, _lib.dptr(_f(A)), A[0].size,
                                _lib.dptr(_f(B)), B[0].size, _lib.dptr(Cf), batch)
    assert rc == 0, _lib.last_error()
    got = np.transpose(Cf, (0, 2, 1))
    assert np.all(np.abs(got - ref32) <= 2 * (k + 2) * 2.0 ** -24 * bound + 1e-30)
    assert np.max(np.abs(got - op[opA](A) @ op[opB](B))) > 1e-9  # really FP32


def test_pagemtimes_broadcast():
    rng = np.random.default_rng(2)
    A = _crand(rng, 1, 40, 40)
    B = _crand(rng, 3, 40, 25)
    Cf = np.zeros((3, 25, 40), complex)
    lib = _lib.load_library()
    rc = lib.orf_pagemtimes(_lib.context(), b"N", b"N", 40, 25, 40, _lib.dptr(_f(A)), 0, _lib.dptr(_f(B)), B[0].size,
                            _lib.dptr(Cf), 3)
    assert rc == 0
    np.testing.assert_allclose(np.transpose(Cf, (0, 2, 1)), A @ B, atol=1e-11)


@pytest.mark.parametrize("n,nrhs,batch", [(5, 3, 4), (61, 122, 3), (98, 294, 2), (242, 242, 2), (300, 7, 2),
                                           (600, 3, 1)])
def test_pagemldivide_and_mrdivide(n, nrhs, batch):
    rng = np.random.default_rng(3)
    A = _crand(rng, batch, n, n) + 2 * np.sqrt(n) * np.eye(n)
    B = _crand(rng, batch, n, nrhs)
    X = np.zeros((batch, nrhs, n), complex)
    lib = _lib.load_library()
    rc = lib.orf_pagemldivide(_lib.context(), n, nrhs, _lib.dptr(_f(A)), n * n, _lib.dptr(_f(B)), _lib.dptr(X), batch)
    assert rc == 0, _lib.last_error()
    np.testing.assert_allclose(np.transpose(X, (0, 2, 1)), np.linalg.solve(A, B), atol=1e-10)
    Br = _crand(rng, batch, nrhs, n)
    Xr = np.zeros((batch, n, nrhs), complex)
    rc = lib.orf_pagemrdivide(_lib.context(), n, nrhs, _lib.dptr(_f(Br)), _lib.dptr(_f(A)), n * n, _lib.dptr(Xr),
                              batch)
    assert rc == 0, _lib.last_error()
    ref = np.swapaxes(np.linalg.solve(np.swapaxes(A, -1, -2), np.swapaxes(Br, -1, -2)), -1, -2)
    np.testing.assert_allclose(np.transpose(Xr, (0, 2, 1)), ref, atol=1e-10)


@pytest.mark.parametrize("n,batch", [(40, 3), (242, 4), (300, 2)])
def test_pagemldivide_general_matrices_pivot(n, batch):
    """No diagonal dominance: nearly every column of the factorization takes an
    interchange (panels of <= 256 rows factor in registers, taller ones in LDS)."""
    rng = np.random.default_rng(11)
    A = _crand(rng, batch, n, n)
    B = _crand(rng, batch, n, 5)
    X = np.zeros((batch, 5, n), complex)
    lib = _lib.load_library()
    rc = lib.orf_pagemldivide(_lib.context(), n, 5, _lib.dptr(_f(A)), n * n, _lib.dptr(_f(B)), _lib.dptr(X), batch)
    assert rc == 0, _lib.last_error()
    ref = np.linalg.solve(A, B)
    np.testing.assert_allclose(np.transpose(X, (0, 2, 1)), ref, atol=1e-10 * np.abs(ref).max())


def test_pagemldivide_needs_pivoting():
    A = np.array([[[0, 1, 2], [3, 0, 1], [1, 1, 0]]], complex)
    B = np.eye(3, dtype=complex)[None]
    X = np.zeros((1, 3, 3), complex)
    lib = _lib.load_library()
    assert lib.orf_pagemldivide(_lib.context(), 3, 3, _lib.dptr(_f(A)), 9, _lib.dptr(_f(B)), _lib.dptr(X), 1) == 0
    np.testing.assert_allclose(np.transpose(X, (0, 2, 1))[0], np.linalg.inv(A[0]), atol=1e-13)


@pytest.mark.parametrize("n,batch", [(6, 3), (25, 4), (61, 3), (61, 520), (98, 2), (131, 2), (242, 1), (339, 1),
                                     (600, 1)])
def test_pageeig(n, batch):
    rng = np.random.default_rng(4)
    A = _crand(rng, batch, n, n)
    w = np.zeros((batch, n), complex)
    V = np.zeros((batch, n, n), complex)
    lib = _lib.load_library()
    rc = lib.orf_pageeig(_lib.context(), n, _lib.dptr(_f(A)), _lib.dptr(w), _lib.dptr(V), batch)
    assert rc == 0, _lib.last_error()
    V = np.transpose(V, (0, 2, 1))
    for q in range(batch):
        res = A[q] @ V[q] - V[q] * w[q][None, :]
        assert np.max(np.abs(res)) < 1e-10 * np.linalg.norm(A[q], 2)
        np.testing.assert_allclose(np.linalg.norm(V[q], axis=0), 1.0, atol=1e-12)
        ref = np.linalg.eigvals(A[q])
        got = w[q]
        # match eigenvalue multisets
        d = np.abs(got[:, None] - ref[None, :])
        assert np.max(np.min(d, axis=1)) < 1e-10 * np.linalg.norm(A[q], 2)


def test_pageeig_degenerate_block_structure():
    """Exactly repeated eigenvalues (as for symmetric pillars) still give a full eigenbasis."""
    rng = np.random.default_rng(5)
    D = np.diag(np.repeat(_crand(rng, 20), 2))
    Q, _ = np.linalg.qr(_crand(rng, 40, 40))
    A = (Q @ D @ np.linalg.inv(Q))[None]
    w = np.zeros((1, 40), complex)
    V = np.zeros((1, 40, 40), complex)
    lib = _lib.load_library()
    assert lib.orf_pageeig(_lib.context(), 40, _lib.dptr(_f(A)), _lib.dptr(w), _lib.dptr(V), 1) == 0
    V = np.transpose(V, (0, 2, 1))[0]
    assert np.max(np.abs(A[0] @ V - V * w[0][None, :])) < 1e-9
    assert np.linalg.cond(V) < 1e8
