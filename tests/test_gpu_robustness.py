"""Failure handling on the device: plan creation that runs out of device memory
half-way (after kernels were queued) returns orf:outOfMemory, hands no block in
use back to the pool, and the solves after it are unchanged."""
import ctypes as C

import numpy as np
import pytest

import opticrcwafdtd_amd as orf
from opticrcwafdtd_amd import _lib
from oracle import rcwa_oracle as O

pytestmark = pytest.mark.gpu
mc = orf.materialConst


def _hip():
    try:
        h = C.CDLL("libamdhip64.so")
    except OSError:
        h = C.CDLL("/opt/rocm/lib/libamdhip64.so")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemGetInfo.argtypes = [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
    return h


def _pillars(radii):
    return [orf.makeStructure([0.4, 0.4], mc(1.45), mc(1),
                              [orf.makeLayer(0.6, mc(1), [orf.shapeCircle(r, mc(3.6 + 0.01j))])]) for r in radii]


def _solve():
    return orf.rcwa2D(_pillars([0.07, 0.13]), 0.94, 0, 0, "p", Nx=4, Nsample=(128, 128), Symmetry="none")


def test_out_of_memory_during_plan_creation_is_clean():
    lib = _lib.load_library()
    _lib.context()
    before = _solve()
    hip = _hip()
    lib.orf_release_cached_memory()
    free, total = C.c_size_t(), C.c_size_t()
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    # leave ~24 GiB: the half-space stage of 3000 wavelengths at +-5 (full basis) needs
    # ~8 GiB per buffer, so its first buffers fit, kernels are queued, and a later one fails
    hog = C.c_void_p()
    keep = 24 << 30
    assert free.value > keep + (8 << 30)
    assert hip.hipMalloc(C.byref(hog), free.value - keep) == 0
    try:
        lams = list(np.linspace(0.90, 0.98, 3000))
        with pytest.raises(orf.OrfError) as e:
            orf.Plan(orf.Problem(_pillars([0.1]), lams, 0, 0, "s", Nx=5, Nsample=(64, 64), Symmetry="none"))
        assert e.value.identifier == "orf:outOfMemory", str(e.value)
    finally:
        hip.hipFree(hog)
        lib.orf_release_cached_memory()
    after = _solve()
    for k in ("R", "T", "r", "t"):
        assert np.array_equal(after[k], before[k])
    o = O.rcwa2d(_pillars([0.07, 0.13]), 0.94, 0, 0, "p", Nx=4, Nsample=(128, 128), Symmetry="none")
    np.testing.assert_allclose(after["T"].T, o["T"], atol=1e-10, rtol=0)
