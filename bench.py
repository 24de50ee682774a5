"""Throughput benchmark of the gfx950 RCWA solve on the metalens phase-library
workload: a-Si pillars (n = 3.6 + 0.01i, h = 0.6 um) in air on SiO2 (lit from
the SiO2 side), square period 0.4 um, 256 radii 40-180 nm x 8 wavelengths
900-980 nm, harmonics +-5 (121 orders, 242 x 242 complex128 eigenproblem),
eps grid 256 x 256, TE and TM: 2048 (cell, wavelength) pages = 4096 solves.

One step = one execution of the resident plan over the whole batch (inputs
already in HBM).  Prints ONE JSON line.  Besides the headline (FP64, full
basis) the line carries, unless ``--quick``: the same batch in the parity-
sector basis (the reference's default Symmetry="auto"), in mixed precision
with its per-efficiency error against FP64, the 4096-cell uneven batch, and
the request loop's p50/p99 latency under Poisson arrivals at three loads.
``--copy-probe`` only runs a device copy of known size (the calibration of
the rocprofv3 FETCH_SIZE / WRITE_SIZE counters).
"""
import argparse
import json
import math
import multiprocessing as mp
import os
import sys
import time

for _v in ("OPENBLAS_NUM_THREADS", "OMP_NUM_THREADS", "MKL_NUM_THREADS"):
    os.environ[_v] = "1"   # one BLAS thread per CPU-baseline worker

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

METRIC = "unit-cell solves/s at ±5 harm. (|ΔT|≤1e-6 vs FP64) + p50/p99 request latency"
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBPS = 8000.0

N_HARM = 5
RADII = np.linspace(0.040, 0.180, 256)
LAMBDAS = np.linspace(0.900, 0.980, 8)


def _progress(msg):
    """Progress on stderr (stdout carries only the JSON line)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def make_devs(orf, radii):
    mc = orf.materialConst
    sio2, air, asi = mc(1.45, "SiO2"), mc(1.0, "air"), mc(3.6 + 0.01j, "a-Si")
    return [orf.makeStructure([0.4, 0.4], sio2, air, [orf.makeLayer(0.6, air, [orf.shapeCircle(r, asi)])])
            for r in radii]


# ---------------------------------------------------------------- CPU baseline (oracle)
def _cpu_worker(args):
    idx, lam_i, sym = args
    radii, lam = [float(RADII[i]) for i in idx], float(LAMBDAS[lam_i])
    import opticrcwafdtd_amd.structure as S
    from oracle import rcwa_oracle as O

    class _NS:
        materialConst = staticmethod(S.materialConst)
        makeStructure = staticmethod(S.makeStructure)
        makeLayer = staticmethod(S.makeLayer)
        shapeCircle = staticmethod(S.shapeCircle)
    devs = make_devs(_NS, radii)
    t0 = time.perf_counter()
    o = O.rcwa2d(devs, lam, 0.0, 0.0, "s", Nx=N_HARM, Nsample=(256, 256), Symmetry=sym)
    return time.perf_counter() - t0, len(radii), (list(idx), lam_i, o["R"], o["T"], o["r"], o["t"])


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(sym, workers, per_worker=12):
    """The FP64 oracle (numpy + LAPACK, one thread per process) on a bounded
    sample of the workload: workers x per_worker TE solves.  Also returns the
    oracle's results, compared with the GPU's pages afterwards."""
    jobs = []
    rng = np.random.default_rng(0)
    for w in range(workers):
        idx = sorted(int(i) for i in rng.choice(len(RADII), size=per_worker, replace=False))
        jobs.append((idx, w % len(LAMBDAS), sym))
    ctx = mp.get_context("fork")
    t0 = time.perf_counter()
    with ctx.Pool(workers) as pool:
        res = pool.map(_cpu_worker, jobs)
    wall = time.perf_counter() - t0
    solves = sum(r[1] for r in res)
    host = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return {"value": solves / wall, "unit": "solves/s", "cores": workers, "kind": "port",
            "cpu_model": _cpu_model(), "host_cpus_visible": host,
            "sample": f"{solves} TE solves (Symmetry={sym}) of the same workload, "
                      f"{per_worker} cells per process, {workers} processes x 1 BLAS thread, "
                      f"{wall:.1f} s wall"}, [r[2] for r in res]


def parity_vs_oracle(samples, gpu):
    """Max |difference| between the oracle's TE results of the CPU sample and the
    GPU's pages of the same (radius, wavelength) in the timed plan."""
    d = {"max_abs_dR": 0.0, "max_abs_dT": 0.0, "max_abs_dr": 0.0, "max_abs_dt": 0.0, "pages": 0}
    for idx, l, R, T, r, t in samples:
        for k, b in enumerate(idx):
            d["max_abs_dR"] = max(d["max_abs_dR"], float(np.max(np.abs(gpu["R"][:, b, l, 0] - R[k]))))
            d["max_abs_dT"] = max(d["max_abs_dT"], float(np.max(np.abs(gpu["T"][:, b, l, 0] - T[k]))))
            d["max_abs_dr"] = max(d["max_abs_dr"], float(np.max(np.abs(gpu["r"][:, :, b, l, 0].T - r[k]))))
            d["max_abs_dt"] = max(d["max_abs_dt"], float(np.max(np.abs(gpu["t"][:, :, b, l, 0].T - t[k]))))
            d["pages"] += 1
    d["within_1e-10"] = max(d["max_abs_dR"], d["max_abs_dT"], d["max_abs_dr"], d["max_abs_dt"]) <= 1e-10
    return d


# ---------------------------------------------------------------- algorithmic flops
def schur_qr_flops(n):
    """Complex QR iteration to Schur form with Schur vectors: Golub & Van Loan's
    25 n^3 (Hessenberg + QR + Q) minus the Hessenberg/Q share (14/3 n^3),
    complex arithmetic counted as 4 real flops."""
    return 4.0 * (25.0 - 14.0 / 3.0) * n ** 3


def solve_flops(n, M):
    """Canonical real flops of one single-patterned-layer solve (the dense-algebra model of the
    standard RCWA chain, independent of the rearrangements the GPU path uses):
    [eps]^-1 8 M^3 + P Q 8 n^3 + eig 100 n^3 + V 8 n^3 + layer S 90 2/3 n^3
    + 2 star products 85 1/3 n^3 each."""
    return 8.0 * M ** 3 + (8 + 100 + 8 + 90 + 2.0 / 3 + 2 * (85 + 1.0 / 3)) * n ** 3


def run(args, sym, precision="fp64", steps=None, warmup=None):
    import opticrcwafdtd_amd as orf
    steps = args.steps if steps is None else steps
    warmup = args.warmup if warmup is None else warmup
    devs = make_devs(orf, RADII)
    prob = orf.Problem(devs, list(LAMBDAS), 0.0, 0.0, ["s", "p"], Nx=N_HARM, Nsample=(256, 256), Symmetry=sym,
                       Precision=precision)
    plan = orf.Plan(prob)
    _progress(f"plan created (Symmetry={sym}, precision={precision}, solved {plan.solved_size})")
    for _ in range(warmup):
        plan.execute()
    _progress("warm-up done")
    plan.enable_timing(True)
    stages = {}  # stage -> [total ms, launches] over every timed step (the plan's timer covers one execution)
    t0 = time.perf_counter()
    for _ in range(steps):
        plan.execute()
        for k, (ms_, cnt) in plan.stage_times().items():
            acc = stages.setdefault(k, [0.0, 0])
            acc[0] += ms_
            acc[1] += cnt
    t1 = time.perf_counter()
    plan.enable_timing(False)
    solves = len(RADII) * len(LAMBDAS) * 2
    _progress(f"{steps} timed steps done")
    out = plan.fetch(amplitudes=True)
    plan.close()
    return {"elapsed": t1 - t0, "steps": steps, "stages": stages, "solves": solves, "n": plan.solved_size,
            "symmetric": plan.symmetric, "pages": plan.solve_pages, "R": out["R"], "T": out["T"],
            "r": out["r"], "t": out["t"]}


def extras(args, full):
    """The other forms of the workload and the other configurations."""
    from opticrcwafdtd_amd import serving, uneven
    ex = {}
    a = run(args, "auto", steps=max(args.steps, 3), warmup=1)
    ex["symmetric_basis"] = {"solves_per_s": round(a["solves"] * a["steps"] / a["elapsed"], 1),
                             "ms_per_step": round(1e3 * a["elapsed"] / a["steps"], 3), "solved_matrix": a["n"],
                             "max_abs_dT_vs_full": float(np.max(np.abs(a["T"] - full["T"])))}
    what = {"mixed": "FP32-MFMA convolution-matrix and star-product GEMMs (operands rounded to FP32, FP32 "
                     "accumulation); eigensolve, propagators and every LU / solve FP64",
            "split": "the same products as hi + lo FP32 split operands (three FP32 MFMA products, lo*lo dropped), "
                     "FP64 accumulation across 4-deep MFMA steps; eigensolve, propagators and LU / solves FP64"}
    for prec in ("mixed", "split"):
        m = run(args, "none", precision=prec, steps=args.steps, warmup=1)
        dT, dR = np.abs(m["T"] - full["T"]), np.abs(m["R"] - full["R"])
        ex["mixed_precision" if prec == "mixed" else "split_precision"] = {
            "solves_per_s": round(m["solves"] * m["steps"] / m["elapsed"], 1),
            "ms_per_step": round(1e3 * m["elapsed"] / m["steps"], 3),
            "speedup_vs_fp64": round((full["elapsed"] / full["steps"]) / (m["elapsed"] / m["steps"]), 4),
            "max_abs_dT": float(dT.max()), "mean_abs_dT": float(dT.mean()),
            "max_abs_dR": float(dR.max()), "mean_abs_dR": float(dR.mean()),
            "within_1e-6": bool(max(dT.max(), dR.max()) <= 1e-6), "what": what[prec]}
    cells = uneven.random_cells(4096, seed=20261203)
    ub = uneven.UnevenBatch(cells, pol=["s", "p"], Nsample=(256, 256), Symmetry="auto")
    t0 = time.perf_counter()
    ub.create()
    t1 = time.perf_counter()
    ub.execute()
    t2 = time.perf_counter()
    ub.execute()
    t3 = time.perf_counter()
    res = ub.results(amplitudes=False)
    t4 = time.perf_counter()
    ub.close()
    ex["uneven_4096"] = {"cells": len(cells), "solves": ub.solves, "size_classes": len(ub.classes), "seed": 20261203,
                         "solves_per_s_execute": round(ub.solves / (t3 - t2), 1),
                         "solves_per_s_end_to_end": round(ub.solves / ((t1 - t0) + (t3 - t2) + (t4 - t3)), 1),
                         "create_s": round(t1 - t0, 3), "execute_s": round(t3 - t2, 3),
                         "max_energy": float(max(float(np.max(r["Rtot"] + r["Ttot"])) for r in res))}
    with serving.Server() as srv:
        serving.run_load(srv, serving.random_requests(40, seed=99), None)  # warm-up
        sat = serving.run_load(srv, serving.random_requests(args.serve_requests, seed=1), None)
        # service capacity under Poisson arrivals: the largest sustained rate, searched from the burst rate
        cap = serving.capacity(srv, sat["req_per_s"], seed=5, seconds=args.serve_seconds)
        C = cap["req_per_s"]
        loads = []
        for f, sd in ((0.5, 2), (0.8, 3), (0.95, 4)):
            n = int(max(args.serve_requests, f * C * args.serve_seconds))
            r = serving.run_load(srv, serving.random_requests(n, seed=sd), f * C, seed=sd)
            loads.append({"load": f, "requests": n, "offered_req_per_s": round(r["offered_req_per_s"], 2),
                          "arrived_req_per_s": round(r["arrived_req_per_s"], 2),
                          "achieved_req_per_s": round(r["steady_req_per_s"], 2),
                          "achieved_frac": round(r["achieved_frac"], 4),
                          "count_frac": round(r["count_frac"], 4),
                          "solves_per_s": round(r["solves_per_s"], 1),
                          "p50_ms": round(r["p50_ms"], 2), "p99_ms": round(r["p99_ms"], 2),
                          "by_kind": r["by_kind"]})
        ex["serving"] = {"mix": "70% lookups +-5, 20% ladders N=3,5,7,9, 10% sidewall sweeps 32 slices x 16 angles; "
                                "TE+TM; Symmetry=auto; seeds 1-5",
                         "burst": {"req_per_s": round(sat["req_per_s"], 2),
                                   "solves_per_s": round(sat["solves_per_s"], 1),
                                   "p50_ms": round(sat["p50_ms"], 2), "p99_ms": round(sat["p99_ms"], 2)},
                         "capacity": {"req_per_s": round(C, 2), "search": cap["search"],
                                      "what": "largest Poisson arrival rate that the service kept up with "
                                              "(achieved_frac >= 0.99 over the same arrival window as the loads, by two request streams), searched from "
                                              "the burst rate: the loads below are fractions of it"},
                         "poisson": loads, "achieved_rate": "over the 10th-90th percentile arrival window: "
                                                            "achieved_frac = service / arrival rate from the "
                                                            "latency trend, 1 / (1 + slope of latency vs "
                                                            "arrival time); count_frac = completions / "
                                                            "arrivals inside the window",
                         "batches": srv.stats["batches"], "max_inflight_batches": srv.stats["max_inflight"]}
    return ex


def measured_traffic(stage, symmetry):
    """HBM bytes per launch of the dominant kernel from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary of this workload (calibrated against a
    copy of known size), or None when no summary covers it."""
    path = os.path.join(ROOT, "profiles", "phase_library_pm5_full_schur_qr_traffic.json")
    if symmetry != "none" or not os.path.exists(path):
        return None
    d = json.load(open(path))
    return d["bytes_per_launch"] if d.get("stage") == stage else None


def copy_probe():
    """One device-to-device copy of 1 GiB, four times (counter calibration)."""
    import torch
    x = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(4):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_per_launch": x.numel() * 8, "launches": 4}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--symmetry", default="none", choices=["none", "auto"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--quick", action="store_true", help="headline only")
    ap.add_argument("--copy-probe", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=300)
    ap.add_argument("--serve-seconds", type=float, default=8.0, help="arrival window per Poisson load level")
    # the GPU host's CPU share is 16 cores per GPU; fewer where fewer are visible
    ap.add_argument("--cpu-workers", type=int,
                    default=min(16, len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity")
                                else (os.cpu_count() or 1)))
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe()
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: every unit cell is an independent solve, so N GPUs would each run "
              "their own share of the batch with no collective on the critical path; only --gpus 1 is "
              "supported.", file=sys.stderr)
        sys.exit(2)
    _progress("CPU baseline (oracle sample)")
    cpu, samples = (None, None) if args.no_cpu else cpu_baseline(args.symmetry, args.cpu_workers)
    _progress("headline plan")
    r = run(args, args.symmetry)
    ms = 1e3 * r["elapsed"] / args.steps
    value = r["solves"] * args.steps / r["elapsed"]
    # dominant kernel over the timed region (HIP events on the plan's stream)
    kernels = {k: v for k, v in r["stages"].items() if k.startswith("eig_")}
    dom = max(kernels, key=lambda k: kernels[k][0])
    tot_ms, launches = kernels[dom]
    n = r["n"]
    pages = r["pages"]
    flops = schur_qr_flops(n) * pages if dom == "eig_schur_qr" else None
    avg_s = tot_ms / launches / 1e3
    achieved = flops / avg_s / 1e12 if flops else None
    M = (2 * N_HARM + 1) ** 2
    traffic = measured_traffic(dom, args.symmetry)
    hbm = None
    if traffic:
        hbm = {"traffic_bytes_per_launch": traffic, "achieved_GBps": round(traffic / avg_s / 1e9, 1),
               "peak_GBps": PEAK_HBM_GBPS, "frac": round(traffic / avg_s / 1e9 / PEAK_HBM_GBPS, 4),
               "what": "measured HBM traffic of the dominant kernel (rocprofv3 FETCH_SIZE / WRITE_SIZE, "
                       "calibrated against a copy) over its launch time: the bandwidth it actually draws"}
    line = {
        "metric": METRIC, "value": round(value, 2), "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "metalens phase library: a-Si pillars r 40-180 nm x 256, lambda 900-980 nm x 8, "
                               "+-5 harmonics, eps 256x256, TE+TM, FP64",
                   "global_batch": r["solves"], "solved_matrix": n, "symmetry": args.symmetry,
                   "pages": pages, "parallelism": "single GPU, batched pages"},
        "roofline": {"bound": "mfma", "kernel": dom, "achieved": round(achieved, 4) if achieved else None,
                     "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / PEAK_FP64_TFLOPS, 5) if achieved else None,
                     "traffic": traffic, "avg_launch_ms": round(tot_ms / launches, 3),
                     "flops_per_launch": flops, "hbm": hbm},
        "whole_solve": {"canonical_flops_per_page": solve_flops(n, M if args.symmetry == "none" else 0),
                        "achieved_tflops": round(solve_flops(n, M if args.symmetry == "none" else 0) *
                                                 pages / (ms / 1e3) / 1e12, 4)},
        "stages_ms_per_step": {k: round(v[0] / r["steps"], 3) for k, v in r["stages"].items()},
        "cpu_baseline": cpu,
        "parity_vs_oracle": parity_vs_oracle(samples, r) if samples else None,
    }
    if not args.quick:
        _progress("extras")
        line["extras"] = extras(args, r)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
