"""The allocation-free left division (orf_pagemldivide_noalloc) through the C
ABI on device buffers: its results equal orf_pagemldivide's on two calls in a
row, the second smaller (the scratch array keeps the first call's contents),
for separate and broadcast A; oversized batches are rejected."""
import ctypes as C

import numpy as np
import pytest

from opticrcwafdtd_amd import _lib

pytestmark = pytest.mark.gpu


class _Hip:
    """The few HIP runtime calls the test needs to own device buffers."""

    def __init__(self):
        try:
            self.h = C.CDLL("libamdhip64.so")
        except OSError:
            self.h = C.CDLL("/opt/rocm/lib/libamdhip64.so")
        self.h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.h.hipFree.argtypes = [C.c_void_p]
        self.h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.h.hipDeviceSynchronize.argtypes = []

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), a.nbytes) == 0
        assert self.h.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return p

    def empty(self, nbytes):
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), nbytes) == 0
        return p

    def download(self, p, like):
        out = np.empty_like(like)
        assert self.h.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0  # device to host
        return out

    def free(self, *ps):
        for p in ps:
            self.h.hipFree(p)


def _system(rng, n, nrhs, batch):
    A = rng.standard_normal((batch, n, n)) + 1j * rng.standard_normal((batch, n, n)) + 4 * n ** 0.5 * np.eye(n)
    B = rng.standard_normal((batch, nrhs, n)) + 1j * rng.standard_normal((batch, nrhs, n))
    # column-major pages: page q of A is A[q].T in memory order
    return np.ascontiguousarray(np.transpose(A, (0, 2, 1))), np.ascontiguousarray(B)


def _reference(lib, ctx, n, nrhs, A, strideA, B, batch):
    X = np.zeros_like(B)
    rc = lib.orf_pagemldivide(ctx, n, nrhs, _lib.dptr(A), strideA, _lib.dptr(B), _lib.dptr(X), batch)
    assert rc == _lib.ORF_OK, _lib.last_error()
    return X


def test_noalloc_equals_allocating_solve_twice():
    lib = _lib.load_library()
    ctx = _lib.context()
    hip = _Hip()
    rng = np.random.default_rng(11)
    for n, nrhs, batch, broadcast in ((96, 40, 24, False), (40, 7, 5, False), (70, 16, 9, True)):
        A, B = _system(rng, n, nrhs, 1 if broadcast else batch)
        if broadcast:
            B = np.ascontiguousarray(rng.standard_normal((batch, nrhs, n)) + 1j * rng.standard_normal((batch, nrhs, n)))
        strideA = 0 if broadcast else 2 * n * n
        ref = _reference(lib, ctx, n, nrhs, A, strideA, B, batch)
        dA, dB = hip.upload(A), hip.upload(B)
        dX = hip.empty(B.nbytes)
        try:
            rc = lib.orf_pagemldivide_noalloc(ctx, n, nrhs, dA, strideA, dB, dX, batch, None)
            assert rc == _lib.ORF_OK, _lib.last_error()
            X = hip.download(dX, B)
            # in place (dX == dB)
            rc = lib.orf_pagemldivide_noalloc(ctx, n, nrhs, dA, strideA, dB, dB, batch, None)
            assert rc == _lib.ORF_OK, _lib.last_error()
            X2 = hip.download(dB, B)
        finally:
            hip.free(dA, dB, dX)
        # same kernels on the same inputs: equal to the last bit (FP64 bar of the
        # library's solves: 1e-10 would do)
        assert np.max(np.abs(X - ref)) == 0.0, np.max(np.abs(X - ref))
        assert np.max(np.abs(X2 - ref)) == 0.0
        # and it solves the system
        Am = np.transpose(A, (0, 2, 1))
        res = np.einsum("bij,bkj->bki", np.broadcast_to(Am, (batch, n, n)), X) - B
        assert np.max(np.abs(res)) < 1e-10 * np.max(np.abs(B)) * n


def test_noalloc_rejects_oversized_batch():
    lib = _lib.load_library()
    ctx = _lib.context()
    dummy = C.c_void_p(1)  # never dereferenced: the size check comes first
    rc = lib.orf_pagemldivide_noalloc(ctx, 768, 1, dummy, 2 * 768 * 768, dummy, dummy, 200, None)
    assert rc == _lib.ORF_INVALID_ARG
    assert "scratch" in _lib.last_error()
    rc = lib.orf_pagemldivide_noalloc(ctx, 8, 1, None, 128, dummy, dummy, 1, None)
    assert rc == _lib.ORF_INVALID_ARG
