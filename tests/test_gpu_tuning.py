"""The eigensolver's tunable kernel choices (orf_tune_*): every feasible choice
of every stage gives the same solve (1e-10 against the built-in choice, which
is itself checked against the oracle elsewhere), forcing and clearing work,
and a solve under the committed tuning table matches the oracle."""
import numpy as np
import pytest

import opticrcwafdtd_amd as orf
from opticrcwafdtd_amd import _lib, autotune
from oracle import rcwa_oracle as O

pytestmark = pytest.mark.gpu
mc = orf.materialConst


def _devs(radii):
    return [orf.makeStructure([0.4, 0.4], mc(1.45), mc(1),
                              [orf.makeLayer(0.6, mc(1), [orf.shapeCircle(r, mc(3.6 + 0.01j))])]) for r in radii]


@pytest.mark.parametrize("N,sym", [(3, "auto"), (5, "auto"), (3, "none"), (5, "none")])
def test_every_kernel_choice_gives_the_same_solve(N, sym):
    lib = _lib.load_library()
    _lib.context()
    devs = _devs([0.07, 0.13])
    ref = orf.rcwa2D(devs, 0.94, 0, 0, ["s", "p"], Nx=N, Nsample=(128, 128), Symmetry=sym)
    n = 2 * (2 * N + 1) ** 2 if sym == "none" else (N + 1) ** 2 + N ** 2
    try:
        for kind in (0, 1, 2):
            for v in autotune.candidates(kind, n):
                assert lib.orf_tune_force(kind, v) == _lib.ORF_OK
                g = orf.rcwa2D(devs, 0.94, 0, 0, ["s", "p"], Nx=N, Nsample=(128, 128), Symmetry=sym)
                for k in ("R", "T", "r", "t"):
                    np.testing.assert_allclose(g[k], ref[k], atol=1e-10, rtol=0, err_msg=f"kind {kind} value {v}")
            assert lib.orf_tune_force(kind, -1) == _lib.ORF_OK
    finally:
        for kind in (0, 1, 2):
            lib.orf_tune_force(kind, -1)


def test_committed_table_solve_matches_oracle():
    lib = _lib.load_library()
    _lib.context()
    import os
    if os.path.exists(_lib.TUNING_TABLE):
        assert lib.orf_tune_load(_lib.TUNING_TABLE.encode()) == _lib.ORF_OK
    devs = _devs([0.09, 0.15])
    g = orf.rcwa2D(devs, 0.92, 0, 0, "s", Nx=4, Nsample=(128, 128), Symmetry="none")
    o = O.rcwa2d(devs, 0.92, 0, 0, "s", Nx=4, Nsample=(128, 128), Symmetry="none")
    np.testing.assert_allclose(g["T"].T, o["T"], atol=1e-10, rtol=0)
    np.testing.assert_allclose(np.transpose(g["t"], (2, 1, 0)), o["t"], atol=1e-10, rtol=0)
