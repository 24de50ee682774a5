"""The request loop: every result is the same alone and inside a mix of
concurrent requests of other sizes, and matches the FP64 oracle."""
import numpy as np
import pytest

import opticrcwafdtd_amd as orf
from opticrcwafdtd_amd import serving as S
from oracle import rcwa_oracle as O

pytestmark = pytest.mark.gpu
TOL = 1e-10


def _small_requests(n, seed):
    reqs = S.random_requests(n, seed, slices=4, angles=3, ladder=(1, 2, 3))
    return [(k, a) for k, a in reqs]


def _submit_small(srv, kind, a):
    if kind == "lookup":
        return srv.submit_lookup(S.pillar(a["r"]), a["lambda0"], N=2)
    if kind == "ladder":
        return srv.submit_ladder(S.pillar(a["r"]), a["lambda0"], Ns=a["Ns"])
    return srv.submit_sweep([S.sidewall_cell(t, r=a["r"], slices=a["slices"]) for t in a["angles"]],
                            a["lambda0"], N=2)


def _same(x, y, tol):
    assert len(x) == len(y)
    for p, q in zip(x, y):
        for k in ("R", "T", "r", "t", "Rtot", "Ttot"):
            np.testing.assert_allclose(p[k], q[k], atol=tol, rtol=0)


def test_alone_equals_in_the_mix_and_oracle():
    probes = [("lookup", {"r": 0.11, "lambda0": 0.93}),
              ("ladder", {"r": 0.07, "lambda0": 0.95, "Ns": (1, 2, 3)}),
              ("sweep", {"r": 0.12, "lambda0": 0.91, "angles": [82.0, 86.0, 90.0], "slices": 4})]
    with S.Server(Nsample=(64, 64), slots=3, memory_budget=8 << 30) as srv:
        alone = []
        for k, a in probes:
            alone.append(_submit_small(srv, k, a).result())
        mix = _small_requests(30, seed=4)
        hs = []
        for i, (k, a) in enumerate(mix):
            hs.append(_submit_small(srv, k, a))
            if i % 10 == 5:
                hs += [_submit_small(srv, k2, a2) for k2, a2 in probes]
        for h in hs:
            h.result()
        assert srv.stats["max_inflight"] >= 2
        # the probes were submitted three times inside the mix
        probe_runs = [hs[i:i + 3] for i in range(len(hs)) if i + 3 <= len(hs) and
                      [x.kind for x in hs[i:i + 3]] == ["lookup", "ladder", "sweep"] and
                      abs(hs[i].items[0].lambda0 - 0.93) < 1e-15]
        assert len(probe_runs) == 3
        for run in probe_runs:
            for h, ref in zip(run, alone):
                _same(h.result(), ref, 1e-12)
    # the lookup alone against the oracle
    o = O.rcwa2d(S.pillar(0.11), 0.93, 0, 0, "s", Nx=2, Nsample=(64, 64))
    np.testing.assert_allclose(alone[0][0]["T"][:, 0], o["T"][0], atol=TOL, rtol=0)
    np.testing.assert_allclose(alone[0][0]["t"][:, :, 0].T, o["t"][0], atol=TOL, rtol=0)
    o = O.rcwa2d(S.sidewall_cell(86.0, r=0.12, slices=4), 0.91, 0, 0, "p", Nx=2, Nsample=(64, 64))
    np.testing.assert_allclose(alone[2][1]["R"][:, 1], o["R"][0], atol=TOL, rtol=0)


def test_memory_budget_limits_inflight_batches():
    with S.Server(Nsample=(64, 64), slots=4, memory_budget=1) as srv:
        hs = [_submit_small(srv, k, a) for k, a in _small_requests(12, seed=9)]
        for h in hs:
            h.result()
        assert srv.stats["max_inflight"] == 1  # nothing fits beside another batch: one at a time
