"""Host-side logic that needs no GPU: explicit-page problems, the uneven-batch
grouping and the seeded mixed-cell workload."""
import math

import numpy as np
import pytest

import opticrcwafdtd_amd as orf
from opticrcwafdtd_amd import uneven

mc = orf.materialConst


def _cell(P, r):
    return orf.makeStructure([P, P], mc(1.45), mc(1), [orf.makeLayer(0.5, mc(1), [orf.shapeCircle(r, mc(2.0))])])


def test_pages_flatten_into_the_problem():
    devs = [_cell(0.4, 0.1), _cell(0.5, 0.12)]
    pr = orf.Problem(devs, [0.9, 1.0, 1.1], 0, 0, "s", Nx=2, Nsample=(64, 64), Pages=[(1, 2), (0, 0)])
    s = pr.struct
    assert s.nPages == 2
    assert [s.pageStruct[i] for i in range(2)] == [1, 0]
    assert [s.pageLambda[i] for i in range(2)] == [2, 0]
    assert [s.structPeriod[i] for i in range(4)] == [0.4, 0.4, 0.5, 0.5]


def test_pages_validation():
    devs = [_cell(0.4, 0.1), _cell(0.5, 0.12)]
    with pytest.raises(orf.OrfError, match="share the period"):
        orf.Problem(devs, 0.9, 0, 0, "s", Nx=2, Nsample=(64, 64))
    with pytest.raises(orf.OrfError, match="out of range"):
        orf.Problem(devs, 0.9, 0, 0, "s", Nx=2, Nsample=(64, 64), Pages=[(2, 0)])
    with pytest.raises(orf.OrfError, match="out of range"):
        orf.Problem(devs, 0.9, 0, 0, "s", Nx=2, Nsample=(64, 64), Pages=[(0, 1)])
    one_d = orf.makeStructure([0.4, math.inf], mc(1), mc(1), [orf.makeLayer(0.1, mc(1))])
    with pytest.raises(orf.OrfError, match="1D or all 2D"):
        orf.Problem([devs[0], one_d], 0.9, 0, 0, "s", Nx=2, Nsample=(64, 64), Pages=[(0, 0)])


def test_mixed_period_default_sampling_is_not_edge_aligned():
    devs = [_cell(0.4, 0.1), _cell(0.5, 0.12)]
    pr = orf.Problem(devs, 0.9, 0, 0, "s", Nx=3, Pages=[(0, 0), (1, 0)])
    assert pr.Nsample == (256, 256)


def test_random_cells_are_seeded_and_cover_the_ranges():
    a = uneven.random_cells(400, seed=11)
    b = uneven.random_cells(400, seed=11)
    assert [c["N"] for c in a] == [c["N"] for c in b]
    assert [c["lambda0"] for c in a] == [c["lambda0"] for c in b]
    Ns = {c["N"] for c in a}
    assert Ns == {3, 4, 5, 6, 7}
    nl = {len(c["dev"]["layers"]) for c in a}
    assert nl == {1, 2, 3, 8}
    P = np.array([c["dev"]["period"][0] for c in a])
    L = np.array([c["lambda0"] for c in a])
    assert P.min() >= 0.3 and P.max() <= 0.6 and L.min() >= 0.45 and L.max() <= 1.6
    kinds = {len(c["dev"]["layers"][0]["shapes"]) for c in a}
    assert 2 in kinds and 1 in kinds  # crosses (two rectangles) and single shapes


def test_grouping_by_size_class():
    cells = uneven.random_cells(200, seed=3)
    g = uneven.group_cells(cells)
    assert sum(len(v) for v in g.values()) == 200
    for (N, nl, _inc, _sub), idx in g.items():
        for i in idx:
            assert cells[i]["N"] == N and len(cells[i]["dev"]["layers"]) == nl
    ub = uneven.UnevenBatch(cells, Nsample=(64, 64))
    assert ub.solves == 400
    work = [ub._work(c["key"], len(c["cells"])) for c in ub.classes]
    assert work == sorted(work, reverse=True)
    for c in ub.classes:
        pr = c["problem"]
        assert pr.struct.nPages == len(c["cells"])


def test_request_mix_is_seeded_with_the_stated_shares():
    from opticrcwafdtd_amd import serving
    a = serving.random_requests(2000, seed=1)
    b = serving.random_requests(2000, seed=1)
    assert a == b
    kinds = [k for k, _ in a]
    for k0 in range(0, len(kinds), 10):  # exact shares in every block of ten
        blk = kinds[k0:k0 + 10]
        assert (blk.count("lookup"), blk.count("ladder"), blk.count("sweep")) == (7, 2, 1)
    assert kinds[:10] != kinds[10:20] or kinds[10:20] != kinds[20:30]  # shuffled within blocks
    sw = next(x for k, x in a if k == "sweep")
    assert len(sw["angles"]) == 16 and sw["slices"] == 32
    assert len(serving.sidewall_cell(85.0, slices=32)["layers"]) == 32
    lad = next(x for k, x in a if k == "ladder")
    assert lad["Ns"] == (3, 5, 7, 9)


# ---------------------------------------------------------------- sloped sidewalls (slantedLayers.m:26-57)
def _halfplane_area(poly, d, n=900):
    """Area of the intersection of the inward-shifted edge half-planes of a convex
    CCW polygon, by sampling (the exact erosion of a convex polygon)."""
    lo, hi = poly.min(axis=0) - 0.01, poly.max(axis=0) + 0.01
    xs, ys = np.linspace(lo[0], hi[0], n), np.linspace(lo[1], hi[1], n)
    X, Y = np.meshgrid(xs, ys)
    ok = np.ones_like(X, bool)
    for i in range(len(poly)):
        p, q = poly[i], poly[(i + 1) % len(poly)]
        t = (q - p) / np.linalg.norm(q - p)
        ok &= (X - p[0]) * -t[1] + (Y - p[1]) * t[0] >= d
    return ok.sum() * (xs[1] - xs[0]) * (ys[1] - ys[0])


def test_slanted_polygon_drops_vanished_edges():
    """A convex pentagon with one short edge: at an inset larger than that edge can take,
    the edge disappears (4 vertices) and the area is that of the shifted half-planes."""
    from opticrcwafdtd_amd.structure import _polygon_offset_inward, _signed_area
    pen = np.array([[0, 0], [0.2, 0], [0.2, 0.1], [0.19, 0.12], [0, 0.15]])
    v = _polygon_offset_inward(pen, 0.03)
    assert v.shape == (4, 2)
    assert abs(_signed_area(v) - _halfplane_area(pen, 0.03)) < 2e-4
    assert _polygon_offset_inward(pen, 0.2) is None   # nothing left


def test_slanted_polygon_split_raises():
    """slantedLayers.m:54-57: a dumbbell whose 0.04-wide neck closes at an inset of 0.02
    raises orf:slantedLayers:split (also rotated); below that it stays one polygon."""
    dumbbell = [[-0.2, -0.1], [-0.05, -0.1], [-0.05, -0.02], [0.05, -0.02], [0.05, -0.1], [0.2, -0.1],
                [0.2, 0.1], [0.05, 0.1], [0.05, 0.02], [-0.05, 0.02], [-0.05, 0.1], [-0.2, 0.1]]
    a = 0.3
    rot = np.array(dumbbell) @ np.array([[np.cos(a), np.sin(a)], [-np.sin(a), np.cos(a)]])
    for poly in (np.array(dumbbell), rot):
        shp = orf.shapePolygon(poly, mc(3.5))
        # 80 deg walls, 0.1 um tall, 4 slices: insets (0.0875 .. 0.0125) * cot(80 deg)
        L = orf.slantedLayers(0.1, mc(1), shp, 80, NSlices=4)
        assert len(L) == 4 and all(len(l["shapes"]) == 1 for l in L)
        with pytest.raises(orf.OrfError) as e:
            orf.slantedLayers(0.3, mc(1), shp, 80, NSlices=4)   # largest inset 0.0463 > 0.02
        assert e.value.identifier == "orf:slantedLayers:split"


def test_capacity_needs_two_streams_to_keep_up(monkeypatch):
    """The capacity search counts a rate as sustained only when two request streams of
    different seeds keep up; a stream that keeps up alone does not raise it."""
    from opticrcwafdtd_amd import serving
    limit = {5: 400.0, 6: 300.0}  # per seed: the rate that stream keeps up with (others: 350)
    seen = []

    def fake_run_load(server, requests, rate, seed=0):
        seen.append((round(rate, 3), seed))
        ok = rate <= limit.get(seed, 350.0)
        return {"achieved_frac": 1.0 if ok else 0.9, "arrived_req_per_s": rate, "steady_req_per_s": rate}

    monkeypatch.setattr(serving, "run_load", fake_run_load)
    monkeypatch.setattr(serving, "random_requests", lambda n, seed: [None] * n)
    cap = serving.capacity(None, 320.0, seed=5, seconds=1.0, steps=6)
    # 320: seed 5 keeps up, seed 6 does not -> not sustained, even though seed 5 alone would be
    assert cap["search"][0]["sustained"] and not cap["search"][1]["sustained"]
    assert cap["req_per_s"] <= 350.0
    assert all(p["sustained"] for p in cap["search"] if p["offered"] <= 300.0)
    # every sustained rate was kept up with by two seeds
    rates = {}
    for p in cap["search"]:
        rates.setdefault(p["offered"], []).append(p["sustained"])
    assert all(len(v) == 2 and all(v) for r, v in rates.items() if r <= cap["req_per_s"])
