"""A request loop over the resident-plan C ABI: continuous batching of unit-cell
solves of differing sizes, admitted by device-memory footprint.

The reference answers every sweep with one blocking ``orf.rcwa2D`` call
(rcwa2D.m:1-36) and its callers, ``buildPhaseLibrary`` (buildPhaseLibrary.m:
1-17) and ``convergeN`` (convergeN.m:20-44), loop over such calls.  Here
requests arrive at any time and are split into solve items (one cell, one
wavelength, one harmonic count, every requested polarisation).  Items of one
size class (harmonic count, layer count, half-spaces, polarisations,
sampling) wait in that class's queue; the scheduler thread takes a class's
whole queue as ONE batch plan (explicit pages: cells of different periods
and wavelengths share launches) when a device slot is free and the plan's
memory estimate fits the budget.  Up to ``slots`` batches run at once, each
on its own HIP stream from its own host thread, so a request that arrives
while others are on the device joins the next batch of its class and runs
beside them.  Each plan keeps its geometry (label grids and their Fourier
transforms), per-context half-space S-matrices and gap modes resident for
the whole batch; the device workspace comes from the library's caching pool
(no hipFree on the request path).

Request kinds (``submit_*``):
  * ``lookup``   one cell at one wavelength (a phase-library entry);
  * ``ladder``   one cell at N = 3, 5, 7, 9 (convergeN's N-ladder, run at once);
  * ``sweep``    a list of cells at one wavelength (e.g. sidewall-angle studies).
Each returns a :class:`Request` handle; ``result()`` waits for it, ``release()``
drops its outputs.
"""
from __future__ import annotations

import collections
import itertools
import threading
import time
from typing import Optional, Sequence

import numpy as np

from .rcwa import Plan, Problem, estimate_memory
from .structure import OrfError


class Request:
    """Handle of one submitted request."""

    _ids = itertools.count()

    def __init__(self, kind: str, items: list):
        self.id = next(Request._ids)
        self.kind = kind
        self.items = items
        self.t_submit = time.perf_counter()
        self.t_done: Optional[float] = None
        self._left = len(items)
        self._results: list = [None] * len(items)
        self._error: Optional[BaseException] = None
        self._event = threading.Event()

    @property
    def solves(self) -> int:
        return sum(len(it.pols) for it in self.items)

    @property
    def latency(self) -> Optional[float]:
        return None if self.t_done is None else self.t_done - self.t_submit

    def done(self) -> bool:
        return self._event.is_set()

    def result(self, timeout=None) -> list:
        """Per item: dict with m, n, R, T (M x nPol), Rtot, Ttot (nPol), r, t (3 x M x nPol), N."""
        if not self._event.wait(timeout):
            raise TimeoutError(f"request {self.id} not finished")
        if self._error is not None:
            raise self._error
        return self._results

    def release(self):
        self._results = [None] * len(self.items)

    def _finish_item(self, k: int, res: dict, t: float):
        self._results[k] = res
        self._left -= 1
        if self._left == 0:
            self.t_done = t
            self._event.set()

    def _fail(self, exc: BaseException):
        self._error = exc
        self.t_done = time.perf_counter()
        self._event.set()


class _Item:
    __slots__ = ("req", "k", "dev", "N", "lambda0", "pols")

    def __init__(self, req, k, dev, N, lambda0, pols):
        self.req, self.k, self.dev, self.N, self.lambda0, self.pols = req, k, dev, N, lambda0, pols


def _mat_key(m: dict):
    return ("const", m["n"]) if m["kind"] == "const" else ("func", id(m))


class Server:
    """Continuous-batching request loop on device 0.

    ``memory_budget``: device bytes the in-flight batches may hold together
    (admission control; a batch whose estimate does not fit waits, and a
    class queue larger than the budget is split).  ``slots``: batches in
    flight at once.  ``max_pages``: cap on the items of one batch.
    """

    def __init__(self, memory_budget: int = 160 << 30, slots: int = 4, max_pages: int = 1024,
                 Nsample=(256, 256), Symmetry="auto", theta=0.0, phi=0.0):
        self.memory_budget = int(memory_budget)
        self.slots = int(slots)
        self.max_pages = int(max_pages)
        self.Nsample = tuple(Nsample)
        self.Symmetry = Symmetry
        self.theta, self.phi = float(theta), float(phi)
        self._queues: dict = collections.OrderedDict()
        self._cv = threading.Condition()
        self._inflight = 0
        self._inflight_bytes = 0
        self._closing = False
        self.stats = {"batches": 0, "pages": 0, "solves": 0, "max_inflight": 0, "max_inflight_bytes": 0}
        self._threads: list = []
        self._sched = threading.Thread(target=self._schedule, name="orf-scheduler", daemon=True)
        self._sched.start()

    # ------------------------------------------------------------ submission
    def _submit(self, kind: str, specs: Sequence[tuple]) -> Request:
        req = Request(kind, [])
        req.items = [_Item(req, k, dev, int(N), float(lm), tuple(pols)) for k, (dev, N, lm, pols) in enumerate(specs)]
        req._left = len(req.items)
        req._results = [None] * len(req.items)
        with self._cv:
            if self._closing:
                raise OrfError("orf:invalidArgument", "server is closed")
            for it in req.items:
                self._queues.setdefault(self._class_key(it), collections.deque()).append(it)
            self._cv.notify_all()
        return req

    def submit_lookup(self, dev: dict, lambda0: float, N: int = 5, pol=("s", "p")) -> Request:
        return self._submit("lookup", [(dev, N, lambda0, pol)])

    def submit_ladder(self, dev: dict, lambda0: float, Ns=(3, 5, 7, 9), pol=("s", "p")) -> Request:
        return self._submit("ladder", [(dev, N, lambda0, pol) for N in Ns])

    def submit_sweep(self, devs: Sequence[dict], lambda0: float, N: int = 5, pol=("s", "p")) -> Request:
        return self._submit("sweep", [(d, N, lambda0, pol) for d in devs])

    def _class_key(self, it: _Item):
        d = it.dev
        return (it.N, len(d["layers"]), _mat_key(d["incident"]), _mat_key(d["substrate"]), it.pols,
                np.isfinite(d["period"][1]))

    # ------------------------------------------------------------ scheduling
    def _problem(self, items: list) -> Problem:
        devs = [it.dev for it in items]
        lams = sorted({it.lambda0 for it in items})
        li = {lm: k for k, lm in enumerate(lams)}
        pages = [(b, li[it.lambda0]) for b, it in enumerate(items)]
        return Problem(devs, lams, self.theta, self.phi, list(items[0].pols), Nx=items[0].N,
                       Nsample=self.Nsample, Symmetry=self.Symmetry, Pages=pages)

    def _estimate(self, prob: Problem, pols) -> int:
        # centred mirror-symmetric cells at normal incidence take the parity-sector basis; once the
        # plan exists its exact footprint replaces this estimate in the in-flight account
        sym = self.Symmetry == "auto" and self.theta == 0.0 and all(p in ("s", "p") for p in pols)
        return estimate_memory(prob, reduced=sym)

    def _pick(self):
        """Oldest class first; returns (key, items, problem, bytes) or None if nothing fits."""
        best = None
        for key, q in self._queues.items():
            if q and (best is None or q[0].req.t_submit < self._queues[best][0].req.t_submit):
                best = key
        if best is None:
            return None
        q = self._queues[best]
        take = min(len(q), self.max_pages)
        while take >= 1:
            items = list(itertools.islice(q, 0, take))
            prob = self._problem(items)
            need = self._estimate(prob, items[0].pols)
            if self._inflight_bytes + need <= self.memory_budget or (self._inflight == 0 and take == 1):
                for _ in range(take):
                    q.popleft()
                return best, items, prob, need
            if self._inflight > 0 and take <= 8:
                return None  # wait for memory to free up rather than shrink further
            take //= 2
        return None

    def _schedule(self):
        while True:
            with self._cv:
                while True:
                    if self._closing and not any(self._queues.values()) and self._inflight == 0:
                        return
                    job = self._pick() if self._inflight < self.slots else None
                    if job is not None:
                        break
                    self._cv.wait(0.05)
                self._inflight += 1
                self._inflight_bytes += job[3]
                self.stats["max_inflight"] = max(self.stats["max_inflight"], self._inflight)
                self.stats["max_inflight_bytes"] = max(self.stats["max_inflight_bytes"], self._inflight_bytes)
            th = threading.Thread(target=self._run_batch, args=job, daemon=True)
            th.start()

    def _run_batch(self, key, items, prob, need):
        plan = None
        try:
            plan = Plan(prob)
            actual = plan.memory()
            with self._cv:
                self._inflight_bytes += actual - need
                need = actual
            plan.execute()
            out = plan.fetch()
            t = time.perf_counter()
            for b, it in enumerate(items):
                res = {"m": out["m"], "n": out["n"], "N": it.N, "symmetry": out["symmetry"],
                       "R": out["R"][:, b], "T": out["T"][:, b], "Rtot": out["Rtot"][b], "Ttot": out["Ttot"][b],
                       "r": out["r"][:, :, b], "t": out["t"][:, :, b]}
                it.req._finish_item(it.k, res, t)
            with self._cv:
                self.stats["batches"] += 1
                self.stats["pages"] += len(items)
                self.stats["solves"] += sum(len(it.pols) for it in items)
        except BaseException as e:  # every request in the batch fails with the library's error
            for it in items:
                if not it.req.done():
                    it.req._fail(e)
        finally:
            if plan is not None:
                plan.close()
            with self._cv:
                self._inflight -= 1
                self._inflight_bytes -= need
                self._cv.notify_all()

    # ------------------------------------------------------------ shutdown
    def drain(self, timeout: Optional[float] = None) -> bool:
        """Wait until every submitted request has finished."""
        t_end = None if timeout is None else time.perf_counter() + timeout
        with self._cv:
            while any(self._queues.values()) or self._inflight > 0:
                left = None if t_end is None else t_end - time.perf_counter()
                if left is not None and left <= 0:
                    return False
                self._cv.wait(0.05 if left is None else min(0.05, left))
        return True

    def close(self):
        with self._cv:
            self._closing = True
            self._cv.notify_all()
        self._sched.join()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


# ---------------------------------------------------------------- the mixed-request workload
def pillar(r: float, n_pillar=3.6 + 0.01j, h=0.6, period=0.4) -> dict:
    """a-Si pillar in air on SiO2, lit from the SiO2 side (the phase-library cell)."""
    from .structure import makeLayer, makeStructure, materialConst as mc, shapeCircle
    return makeStructure([period, period], mc(1.45), mc(1), [makeLayer(h, mc(1), [shapeCircle(r, mc(n_pillar))])])


def sidewall_cell(angle: float, r=0.12, h=0.6, slices=32, period=0.4) -> dict:
    from .structure import makeStructure, materialConst as mc, shapeCircle, slantedLayers
    return makeStructure([period, period], mc(1.45), mc(1),
                         slantedLayers(h, mc(1), shapeCircle(r, mc(3.6 + 0.01j)), angle, NSlices=slices))


def random_requests(count: int, seed: int, slices=32, angles=16, ladder=(3, 5, 7, 9)) -> list:
    """Seeded request mix: 70 % single-wavelength +-5 lookups, 20 % N-ladders
    3-5-7-9 on one geometry, 10 % sidewall sweeps (``slices`` slices x
    ``angles`` angles).  The proportions hold exactly in every block of ten
    requests (seven lookups, two ladders, one sweep, in a shuffled order), so
    that the work offered per second does not vary with the seed: a sweep is
    ``slices`` x ``angles`` x 2 solves, about 95 % of the mix's work.  Each
    entry is (kind, args) for :func:`submit`."""
    rng = np.random.Generator(np.random.PCG64(seed))
    block = ["lookup"] * 7 + ["ladder"] * 2 + ["sweep"]
    out = []
    for k in range(count):
        if k % 10 == 0:
            kinds = list(rng.permutation(block))
        kind = kinds[k % 10]
        lam = float(rng.uniform(0.90, 0.98))
        if kind == "lookup":
            out.append(("lookup", {"r": float(rng.uniform(0.04, 0.18)), "lambda0": lam}))
        elif kind == "ladder":
            out.append(("ladder", {"r": float(rng.uniform(0.04, 0.18)), "lambda0": lam, "Ns": tuple(ladder)}))
        else:
            r = float(rng.uniform(0.08, 0.16))
            out.append(("sweep", {"r": r, "lambda0": lam,
                                  "angles": [float(a) for a in np.linspace(80.0, 90.0, angles)],
                                  "slices": slices}))
    return out


def submit(server: Server, kind: str, a: dict) -> Request:
    if kind == "lookup":
        return server.submit_lookup(pillar(a["r"]), a["lambda0"])
    if kind == "ladder":
        return server.submit_ladder(pillar(a["r"]), a["lambda0"], Ns=a["Ns"])
    if kind == "sweep":
        return server.submit_sweep([sidewall_cell(t, r=a["r"], slices=a["slices"]) for t in a["angles"]], a["lambda0"])
    raise OrfError("orf:invalidArgument", f"unknown request kind {kind}")


def run_load(server: Server, requests: Sequence[tuple], rate: Optional[float], seed: int = 0) -> dict:
    """Submit ``requests`` with Poisson arrivals at ``rate`` per second (all at
    once when ``rate`` is None), wait for all, and report latency percentiles
    and throughput.

    ``req_per_s`` / ``solves_per_s`` count requests between the first arrival
    and the last completion (they include the drain tail).  The steady state
    is read over the middle 80 % of the arrivals (10th to 90th percentile
    arrival time), where the queue is neither filling from empty nor
    draining:

    * ``achieved_frac`` = service rate / arrival rate from the latency trend:
      a request arriving at t waits L(t); when arrivals come at rate a and
      completions at rate c < a, the backlog grows and L rises with slope
      a / c - 1, so c / a = 1 / (1 + slope), the slope being the least-squares
      line through (arrival time, latency) of the window's requests.  A
      sustained load has slope ~0 and ``achieved_frac`` ~1.  It uses every
      request of the window, so it is far less noisy than counts at the
      window's two ends, which move by the backlog present at those instants;
    * ``arrived_req_per_s``: arrivals in the window / its length (the offered
      rate as realised, Poisson scatter included); ``steady_req_per_s`` =
      ``arrived_req_per_s`` x min(1, ``achieved_frac``);
    * ``count_frac``: completions / arrivals inside the window, for reference.
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    gaps = rng.exponential(1.0 / rate, len(requests)) if rate else np.zeros(len(requests))
    handles = []
    t0 = time.perf_counter()
    t_next = t0
    for (kind, a), g in zip(requests, gaps):
        t_next += g
        d = t_next - time.perf_counter()
        if d > 0:
            time.sleep(d)
        handles.append(submit(server, kind, a))
    for h in handles:
        h.result()
    t_end = max(h.t_done for h in handles)
    lat = np.array([h.latency for h in handles])
    solves = sum(h.solves for h in handles)
    arrivals = np.array([h.t_submit for h in handles])
    done = np.array([h.t_done for h in handles])
    a_lo, a_hi = np.percentile(arrivals, 10), np.percentile(arrivals, 90)
    window = a_hi - a_lo
    n_done = int(np.sum((done >= a_lo) & (done < a_hi)))
    sel = (arrivals >= a_lo) & (arrivals < a_hi)
    n_arr = int(np.sum(sel))
    arrived = n_arr / window if window > 0 else float("nan")
    frac = None
    if rate and n_arr > 2:
        slope = float(np.polyfit(arrivals[sel] - a_lo, lat[sel], 1)[0])
        frac = 1.0 / (1.0 + slope) if slope > -0.5 else 2.0
    steady = arrived * min(1.0, frac) if frac is not None else float("nan")
    by_kind = {}
    for k in sorted({h.kind for h in handles}):
        lk = np.array([h.latency for h in handles if h.kind == k])
        by_kind[k] = {"count": int(lk.size), "p50_ms": float(np.percentile(lk, 50) * 1e3),
                      "p99_ms": float(np.percentile(lk, 99) * 1e3)}
    return {"offered_req_per_s": rate, "requests": len(handles), "solves": solves,
            "elapsed_s": t_end - t0, "solves_per_s": solves / (t_end - t0),
            "req_per_s": len(handles) / (t_end - t0), "steady_req_per_s": steady,
            "arrived_req_per_s": arrived, "achieved_frac": frac,
            "count_frac": n_done / n_arr if rate and n_arr else None,
            "p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
            "by_kind": by_kind}


def capacity(server: Server, rate0: float, seed: int, seconds: float = 5.0, steps: int = 6,
             sustain: float = 0.99) -> dict:
    """Largest sustained Poisson arrival rate: starting at ``rate0`` (e.g. the
    burst rate), each step offers ``seconds`` worth of requests and raises the
    rate by 15 % while the service keeps up with the arrivals over the
    steady window (``achieved_frac`` of :func:`run_load` >= ``sustain``), else
    bisects towards the last sustained rate.  A rate counts as sustained only
    when a second request stream (another seed) keeps up too: at one offered
    rate the achieved rate of single streams scatters by 5-10 % (arrival
    sequence, request parameters, host timing), so one lucky stream would set
    the capacity too high.  Returns the sustained rate and the probes taken."""
    lo, hi, rate, log = None, None, float(rate0), []
    for k in range(steps):
        ok = True
        for rep in range(2):
            sd = seed + 2 * k + rep
            n = int(max(200, rate * seconds))
            r = run_load(server, random_requests(n, sd), rate, seed=sd)
            ok_r = r["achieved_frac"] >= sustain
            log.append({"offered": round(rate, 2), "seed": sd, "arrived": round(r["arrived_req_per_s"], 2),
                        "achieved": round(r["steady_req_per_s"], 2), "sustained": bool(ok_r)})
            if not ok_r:
                ok = False
                break
        if ok:
            lo = rate
            rate = rate * 1.15 if hi is None else 0.5 * (rate + hi)
        else:
            hi = rate
            rate = 0.85 * rate if lo is None else 0.5 * (lo + rate)
    return {"req_per_s": lo if lo is not None else rate, "search": log}


__all__ = ["Server", "Request", "random_requests", "submit", "run_load", "capacity", "pillar", "sidewall_cell"]
