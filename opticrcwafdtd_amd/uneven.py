"""Uneven unit-cell batches: cells that differ in harmonic count, layer count,
period and wavelength solved together on one device.

The reference solves one batch per ``orf.rcwa2D`` call and requires the
batch to share the period, the half-spaces and the layer count
(rcwa2D.m:4-7); a sweep over mixed cells is a loop of such calls.  Here the
cells are grouped into size classes (harmonic count, layer count, half-space
materials); each class becomes ONE resident plan whose pages are explicit
(structure, wavelength) pairs, so cells of different periods and
wavelengths share the same launches (each page carries its own k-vectors
and half-space S-matrices).  The classes run concurrently, one plan per
host thread and one HIP stream per plan, so small classes fill the device
beside large ones.
"""
from __future__ import annotations

import threading
from concurrent.futures import ThreadPoolExecutor
from typing import Sequence

import numpy as np

from .rcwa import Plan, Problem
from .structure import (OrfError, makeLayer, makeStructure, materialConst, shapeCircle, shapeCross,
                        shapeRect, slantedLayers)


def _half_space_key(m: dict):
    return ("const", m["n"]) if m["kind"] == "const" else ("func", id(m))


def group_cells(cells: Sequence[dict]) -> dict:
    """Size classes: (N, nLayers, incident, substrate) -> list of cell indices."""
    groups: dict = {}
    for i, c in enumerate(cells):
        d = c["dev"]
        key = (int(c["N"]), len(d["layers"]), _half_space_key(d["incident"]), _half_space_key(d["substrate"]))
        groups.setdefault(key, []).append(i)
    return groups


class UnevenBatch:
    """Resident plans for a list of cells ``{"dev": structure, "N": int, "lambda0": float}``.

    ``pol`` is one polarisation or a list (TE and TM are two solves of one
    cell).  ``execute`` runs every class; ``results`` returns one dict per
    cell with the reference's output fields (``m, n, R, T, Rtot, Ttot, r,
    t``; R is M[, nPol]).
    """

    def __init__(self, cells: Sequence[dict], pol=("s", "p"), theta=0.0, phi=0.0, Nsample=(256, 256),
                 Symmetry="auto", Factorization="li", workers=4):
        if not cells:
            raise OrfError("orf:invalidArgument", "no cells")
        self.cells = list(cells)
        self.pol = list(pol) if isinstance(pol, (list, tuple)) else pol
        self.workers = max(1, int(workers))
        self.classes = []
        for key, idx in sorted(group_cells(self.cells).items(), key=lambda kv: -self._work(kv[0], len(kv[1]))):
            devs = [self.cells[i]["dev"] for i in idx]
            lams = sorted({float(self.cells[i]["lambda0"]) for i in idx})
            li = {lm: k for k, lm in enumerate(lams)}
            pages = [(b, li[float(self.cells[i]["lambda0"])]) for b, i in enumerate(idx)]
            N = key[0]
            prob = Problem(devs, lams, theta, phi, self.pol, Nx=N, Nsample=Nsample, Symmetry=Symmetry,
                           Factorization=Factorization, Pages=pages)
            self.classes.append({"key": key, "cells": idx, "problem": prob, "plan": None})
        self._lock = threading.Lock()
        self.graphs = False

    @staticmethod
    def _work(key, count):
        N, nl = key[0], key[1]
        return count * nl * (2 * (2 * N + 1) ** 2) ** 3

    @property
    def solves(self) -> int:
        npol = len(self.pol) if isinstance(self.pol, list) else 1
        return len(self.cells) * npol

    def _run_class(self, c):
        if c["plan"] is None:
            c["plan"] = Plan(c["problem"])
            if self.graphs:
                c["plan"].enable_graphs(True)
        c["plan"].execute()

    def enable_graphs(self, on=True):
        """Capture every class plan's stage chain into HIP graphs and replay them
        (orf_plan_enable_graphs) from the next execute on."""
        self.graphs = bool(on)
        for c in self.classes:
            if c["plan"] is not None:
                c["plan"].enable_graphs(self.graphs)

    def create(self):
        """Create every plan (uploads, rasterisation, half-space S-matrices, workspace)."""
        for c in self.classes:
            if c["plan"] is None:
                c["plan"] = Plan(c["problem"])
                if self.graphs:
                    c["plan"].enable_graphs(True)

    def execute(self):
        """Run all classes, ``workers`` at a time, largest first."""
        with ThreadPoolExecutor(self.workers) as ex:
            for f in [ex.submit(self._run_class, c) for c in self.classes]:
                f.result()

    def results(self, amplitudes=True) -> list:
        out = [None] * len(self.cells)
        for c in self.classes:
            o = c["plan"].fetch(amplitudes=amplitudes)
            for b, i in enumerate(c["cells"]):
                cell = {"m": o["m"], "n": o["n"], "symmetry": o["symmetry"], "N": c["key"][0]}
                for k in ("R", "T"):
                    cell[k] = o[k][:, b]
                for k in ("Rtot", "Ttot"):
                    cell[k] = o[k][b]
                if amplitudes:
                    for k in ("r", "t"):
                        cell[k] = o[k][:, :, b]
                out[i] = cell
        return out

    def close(self):
        for c in self.classes:
            if c["plan"] is not None:
                c["plan"].close()
                c["plan"] = None


def solveUneven(cells, pol=("s", "p"), theta=0.0, phi=0.0, Nsample=(256, 256), Symmetry="auto",
                workers=4, amplitudes=True) -> list:
    """One call over an uneven list of cells; returns one output dict per cell."""
    ub = UnevenBatch(cells, pol=pol, theta=theta, phi=phi, Nsample=Nsample, Symmetry=Symmetry,
                     workers=workers)
    try:
        ub.execute()
        return ub.results(amplitudes=amplitudes)
    finally:
        ub.close()


# ---------------------------------------------------------------- the mixed-cell workload
MATERIALS = {"TiO2": 2.5, "Si3N4": 2.0, "a-Si": 3.6 + 0.01j, "SiO2": 1.45}


def random_cells(count: int, seed: int, N_range=(3, 7), period_um=(0.30, 0.60), lambda_um=(0.45, 1.60),
                 slices=8) -> list:
    """Seeded mixed unit cells: circle, rectangle (sides 60-300 nm) or cross
    (arm width 40-120 nm) pillars of TiO2 / Si3N4 / a-Si in air on SiO2 (lit
    from the SiO2 side); 1-3 stacked uniform layers or one sloped-sidewall
    pillar (80-90 deg) staircased into ``slices`` layers; harmonics
    +-N_range, period and wavelength uniform in the given ranges."""
    rng = np.random.Generator(np.random.PCG64(seed))
    air, sio2 = materialConst(1.0), materialConst(MATERIALS["SiO2"])
    pillar_mats = [materialConst(MATERIALS[k], k) for k in ("TiO2", "Si3N4", "a-Si")]
    cells = []
    for _ in range(count):
        N = int(rng.integers(N_range[0], N_range[1] + 1))
        period = float(rng.uniform(*period_um))
        lam = float(rng.uniform(*lambda_um))
        kind = int(rng.integers(0, 3))

        def shape(mat):
            if kind == 0:
                r = float(rng.uniform(0.03, 0.45 * period))
                return [shapeCircle(r, mat)]
            if kind == 1:
                wx = float(rng.uniform(0.06, min(0.30, 0.95 * period)))
                wy = float(rng.uniform(0.06, min(0.30, 0.95 * period)))
                return [shapeRect(wx, wy, mat)]
            w = float(rng.uniform(0.04, 0.12))
            arm = float(rng.uniform(max(w + 0.02, 0.1), 0.95 * period))
            return shapeCross(w, arm, mat)

        if rng.uniform() < 0.5:
            nl = int(rng.integers(1, 4))
            layers = []
            for _j in range(nl):
                mat = pillar_mats[int(rng.integers(0, 3))]
                layers.append(makeLayer(float(rng.uniform(0.1, 0.4)), air, shape(mat)))
        else:
            mat = pillar_mats[int(rng.integers(0, 3))]
            angle = float(rng.uniform(80.0, 90.0))
            layers = slantedLayers(float(rng.uniform(0.3, 0.8)), air, shape(mat), angle, NSlices=slices)
        dev = makeStructure([period, period], sio2, air, layers)
        cells.append({"dev": dev, "N": N, "lambda0": lam})
    return cells


def cell_flops(cell: dict) -> float:
    """Canonical FP64 flop count of one cell's solve (per polarisation) at full basis."""
    N = int(cell["N"])
    n = 2 * (2 * N + 1) ** 2
    M = n // 2
    nl = len(cell["dev"]["layers"])
    patterned = sum(1 for L in cell["dev"]["layers"] if L["shapes"])
    per_layer = 8 * n ** 3 + 100 * n ** 3 + 8 * n ** 3 + 90.67 * n ** 3 + 85.33 * n ** 3
    return patterned * (8 * M ** 3) + nl * per_layer + 85.33 * n ** 3


__all__ = ["UnevenBatch", "solveUneven", "random_cells", "group_cells", "cell_flops", "MATERIALS"]
