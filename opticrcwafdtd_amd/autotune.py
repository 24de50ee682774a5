"""Autotuner of the batched eigensolver's kernel choices.

For each solved matrix size that the library's workloads produce (parity-
sector bases at +-2..+-9 harmonics, full bases at +-2..+-6) and each batch
class (a few pages / >= 512 pages), it times every feasible choice of the
three eigensolver stages on a real plan (a-Si pillar cells, 256 x 256
permittivity grid, the plan's own stage timers):

* Hessenberg reduction: rank-1 updates (0) or 8 / 12-wide panels;
* Schur QR: per-rotation kernel with H in LDS (0, n <= 101) or multishift
  passes with early deflation, 4 bulges and a 12-wide window (1) or 8 bulges
  and a 16-wide window (2), each also in two-wave workgroups with H and Z in
  global memory, six / four per CU (3 / 4);
* eigenvectors: 16 / 32 rows per lane (n <= 256 / 512) or one wave per
  vector (0);

and writes the fastest to ``tuning_table.txt`` beside this file
("kind n large_batch value" per line), which the host loads at context
creation (orf_tune_load).  Run on the target GPU:

    python -m opticrcwafdtd_amd.autotune [--out PATH] [--quick] [--kinds 0,1,2]

``--kinds`` re-tunes only the listed kinds and keeps the other kinds' rows of
the existing table.
"""
from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np

from . import _lib
from .rcwa import Plan, Problem
from .structure import makeLayer, makeStructure, materialConst, shapeCircle

KINDS = {0: ("eig_schur_qr", "qr"), 1: ("eig_hessenberg", "hessenberg"), 2: ("eig_vectors", "vectors")}
LDS_BYTES = 163840          # a workgroup's LDS (160 KiB); the per-rotation QR keeps H there


def candidates(kind: int, n: int) -> list:
    if kind == 0:
        return ([0] if n * n * 16 + 64 <= LDS_BYTES else []) + [1, 3, 2, 4]
    if kind == 1:
        return [0] + ([12, 8] if n >= 32 else [])
    return ([16] if n <= 256 else []) + ([32] if n <= 512 else []) + [0]


def _cells(radii):
    mc = materialConst
    return [makeStructure([0.4, 0.4], mc(1.45), mc(1), [makeLayer(0.6, mc(1), [shapeCircle(r, mc(3.6 + 0.01j))])])
            for r in radii]


def _plan(N: int, symmetry: str, pages: int) -> Plan:
    nl = 8 if pages >= 64 else 1
    radii = np.linspace(0.04, 0.18, pages // nl)
    lams = list(np.linspace(0.90, 0.98, nl)) if nl > 1 else 0.94
    return Plan(Problem(_cells(radii), lams, 0.0, 0.0, ["s", "p"], Nx=N, Nsample=(256, 256), Symmetry=symmetry))


def _stage_ms(plan: Plan, stage: str, reps: int) -> float:
    plan.enable_timing(True)
    best = float("inf")
    for _ in range(reps):
        plan.execute()
        best = min(best, plan.stage_times().get(stage, (float("inf"), 1))[0])
    plan.enable_timing(False)
    return best


def tune(sizes, batches, reps=2, log=print, kinds=tuple(KINDS)) -> list:
    lib = _lib.load_library()
    _lib.context()
    lib.orf_tune_clear()
    rows = []
    for N, symmetry in sizes:
        for large, pages in batches:
            plan = _plan(N, symmetry, pages)
            n = plan.solved_size
            plan.execute()  # warm-up
            for kind, (stage, label) in KINDS.items():
                if kind not in kinds:
                    continue
                times = {}
                for v in candidates(kind, n):
                    lib.orf_tune_force(kind, v)
                    times[v] = _stage_ms(plan, stage, reps)
                lib.orf_tune_force(kind, -1)
                best = min(times, key=times.get)
                rows.append((kind, n, large, best, times))
                log(f"n={n:4d} ({symmetry:4s} N={N}) pages={pages:5d} {label:10s} " +
                    " ".join(f"{v}:{t:8.3f}ms" for v, t in times.items()) + f"  -> {best}")
            plan.close()
    return rows


def write_table(rows, path):
    with open(path, "w") as f:
        f.write("# kind n large_batch value -- written by opticrcwafdtd_amd/autotune.py\n")
        f.write("# kinds: 0 Schur QR (0 per-rotation, 1 multishift 4 bulges, 2 multishift 8 bulges,\n")
        f.write("#        3 / 4 multishift 4 / 8 bulges in two-wave workgroups);\n")
        f.write("#        1 Hessenberg (0 rank-1, 8 / 12 panel);\n")
        f.write("#        2 eigenvectors (16 / 32 rows per lane, 0 one wave per vector)\n")
        for kind, n, large, best, times in sorted(rows, key=lambda r: (r[0], r[2], r[1])):
            note = times if isinstance(times, str) else " ".join(f"{v}:{t:.3f}" for v, t in times.items())
            f.write(f"{kind} {n} {large} {best}   # {note}\n")


def read_table(path) -> list:
    """Rows of an existing table as (kind, n, large, value, note)."""
    rows = []
    if os.path.exists(path):
        for line in open(path):
            body, _, note = line.partition("#")
            f = body.split()
            if len(f) == 4:
                rows.append((int(f[0]), int(f[1]), int(f[2]), int(f[3]), note.strip()))
    return rows


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--out", default=_lib.TUNING_TABLE)
    ap.add_argument("--quick", action="store_true", help="two sizes, small batches only")
    ap.add_argument("--kinds", default="0,1,2", help="kinds to tune (others are kept from --out)")
    args = ap.parse_args(argv)
    kinds = tuple(int(k) for k in args.kinds.split(","))
    if args.quick:
        sizes, batches = [(3, "auto"), (3, "none")], [(0, 8)]
    else:
        sizes = [(N, "auto") for N in (2, 3, 4, 5, 6, 7, 9)] + [(N, "none") for N in (2, 3, 4, 5, 6)]
        batches = [(0, 8), (1, 1024)]
    t0 = time.perf_counter()
    rows = tune(sizes, batches, log=lambda s: print(s, flush=True), kinds=kinds)
    rows += [r for r in read_table(args.out) if r[0] not in kinds]
    write_table(rows, args.out)
    print(f"wrote {args.out} ({len(rows)} entries, {time.perf_counter() - t0:.0f} s)", file=sys.stderr)


if __name__ == "__main__":
    main()
