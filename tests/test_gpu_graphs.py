"""Graph-captured execution (orf_plan_enable_graphs): every replay of a plan's
captured chunks gives the eager results to the last bit, with other plans'
eager work and host synchronisations between the replays; the eager results
themselves match the oracle."""
import numpy as np
import pytest

import opticrcwafdtd_amd as orf
from opticrcwafdtd_amd import serving, uneven
from oracle import rcwa_oracle as O

pytestmark = pytest.mark.gpu
mc = orf.materialConst


def _pillars(radii, h=0.6):
    return [orf.makeStructure([0.4, 0.4], mc(1.45), mc(1),
                              [orf.makeLayer(h, mc(1), [orf.shapeCircle(r, mc(3.6 + 0.01j))])]) for r in radii]


def _stack(r):
    si = mc(3.5)
    return orf.makeStructure([0.5, 0.5], mc(1), mc(1.45),
                             [orf.makeLayer(0.2, mc(1), [orf.shapeCircle(r, si)]), orf.makeLayer(0.05, mc(2.0)),
                              orf.makeLayer(0.1, mc(1), [orf.shapeRect(0.3, 0.2, si)])])


CASES = {
    "lookup_sector": lambda: orf.Problem(_pillars([0.11]), 0.93, 0, 0, ["s", "p"], Nx=5, Nsample=(128, 128)),
    "full_basis_batch": lambda: orf.Problem(_pillars(np.linspace(0.05, 0.17, 24)), [0.9, 0.95], 0, 0, ["s", "p"],
                                            Nx=4, Nsample=(128, 128), Symmetry="none"),
    "sidewall_layers": lambda: orf.Problem([serving.sidewall_cell(a, slices=8) for a in (82.0, 86.0, 90.0)], 0.91,
                                           0, 0, ["s", "p"], Nx=3, Nsample=(128, 128)),
    "chunked_stack_conical": lambda: orf.Problem([_stack(r) for r in (0.08, 0.11, 0.14, 0.17, 0.2)], 1.0, 0.2, 0.3,
                                                 "p", Nx=3, Nsample=(96, 96), BatchSize=2),
}


def _fetch(plan):
    o = plan.fetch()
    return {k: np.array(o[k]) for k in ("R", "T", "Rtot", "Ttot", "r", "t")}


@pytest.mark.parametrize("case", sorted(CASES))
def test_graph_replays_equal_eager(case):
    plan = orf.Plan(CASES[case]())
    other = orf.Plan(CASES["lookup_sector"]())
    try:
        plan.execute()
        eager = _fetch(plan)
        plan.enable_graphs(True)
        for rep in range(4):          # capture on the first, replay on the others
            plan.execute()
            got = _fetch(plan)
            for k in eager:
                assert np.array_equal(got[k], eager[k]), (case, rep, k)
            other.execute()           # eager work of another plan between replays
            _fetch(other)
        plan.enable_graphs(False)
        plan.execute()
        got = _fetch(plan)
        for k in eager:
            assert np.array_equal(got[k], eager[k]), (case, "eager again", k)
    finally:
        plan.close()
        other.close()


def test_graph_lookup_matches_oracle():
    plan = orf.Plan(CASES["lookup_sector"]())
    try:
        plan.enable_graphs(True)
        for _ in range(3):
            plan.execute()
        g = plan.fetch()
    finally:
        plan.close()
    o = O.rcwa2d(_pillars([0.11]), 0.93, 0, 0, "p", Nx=5, Nsample=(128, 128))
    np.testing.assert_allclose(g["T"][:, 0, 1], o["T"][0], atol=1e-10, rtol=0)
    np.testing.assert_allclose(g["t"][:, :, 0, 1].T, o["t"][0], atol=1e-10, rtol=0)


def test_uneven_batch_with_graphs_equals_eager():
    cells = uneven.random_cells(160, seed=5, N_range=(2, 4), slices=4)
    ub = uneven.UnevenBatch(cells, pol=["s", "p"], Nsample=(96, 96))
    try:
        ub.execute()
        eager = ub.results()
        ub.enable_graphs(True)
        for _ in range(3):
            ub.execute()
            got = ub.results()
            for a, b in zip(got, eager):
                for k in ("R", "T", "r", "t"):
                    assert np.array_equal(a[k], b[k])
    finally:
        ub.close()
