"""Python host mirror of the reference's RCWA API over the gfx950 C ABI.

``rcwa2D`` keeps the signature, option names, output fields and error
identifiers of ``orf.rcwa2D`` (toolbox/+orf/rcwa2D.m:1-35); ``rcwa1D``,
``buildPhaseLibrary`` and ``convergeN`` mirror rcwa1D.m:1-28,
buildPhaseLibrary.m:1-17 and convergeN.m:1-45.  Every solve runs on the GPU
through liborf_rcwa.so; this module only flattens the device descriptions.

Output arrays follow MATLAB's order: ``R``, ``T`` are M x B, ``Rtot``,
``Ttot`` are B, ``r``, ``t`` are 3 x M x B.  Two build extensions append
trailing dimensions: a vector ``lambda0`` (wavelength axis) and a list of
polarisations (polarisation axis).  A third, ``Pages``, replaces the
(structure, wavelength) product by an explicit list of pairs and lets the
structures differ in period; the outputs are then M x nPages.
"""
from __future__ import annotations

import ctypes as C
import math
from typing import Callable, Optional, Sequence

import numpy as np

from . import _lib
from .structure import OrfError, OrfWarning, alignedSampleCount, default_sample_counts, layerEps, material_index

_SHAPE_TYPES = {"circle": 0, "rect": 1, "polygon": 2}


def _pol_kinds(pol):
    """Normalise pol to a list of (kind, aTE, aTM) and whether it was a list."""
    def one(p):
        if isinstance(p, str):
            if p not in ("s", "p"):
                raise OrfError("rcwa:pol", 'pol must be "s", "p" or [aTE aTM].')
            return (0 if p == "s" else 1, 0j, 0j)
        j = np.asarray(p, dtype=complex).ravel()
        if j.size != 2:
            raise OrfError("rcwa:pol", "Jones vector must be [aTE aTM].")
        return (2, complex(j[0]), complex(j[1]))
    if isinstance(pol, (list, tuple)) and len(pol) > 0 and (
            isinstance(pol[0], str) or (not np.isscalar(pol[0]) and np.asarray(pol[0]).size == 2)):
        return [one(p) for p in pol], True
    return [one(pol)], False


class Problem:
    """Flattened orf_problem with the numpy buffers it points to kept alive."""

    def __init__(self, devs, lambda0, theta, phi, pol, Nx=5, Ny=None, Nsample=0,
                 Factorization="li", GapEps=float("nan"), Symmetry="auto", BatchSize=0,
                 Precision="fp64", Pages=None):
        if isinstance(devs, dict):
            devs = [devs]
        devs = list(devs)
        if not devs:
            raise OrfError("orf:invalidArgument", "devs is empty")
        self.devs = devs
        Lx, Ly = devs[0]["period"]
        self.pages = None if Pages is None else np.asarray(Pages, np.int32).reshape(-1, 2)
        periods = np.array([list(d["period"]) for d in devs], float)
        for d in devs:
            if list(d["period"]) != [Lx, Ly] and self.pages is None:
                raise OrfError("orf:invalidArgument", "structures must share the period")
            if math.isfinite(d["period"][1]) != math.isfinite(Ly):
                raise OrfError("orf:invalidArgument", "structures must all be 1D or all 2D")
        if Ny is None:
            Ny = Nx
        if not math.isfinite(Ly):
            Ny = 0
        if Factorization not in ("li", "laurent"):
            raise OrfError("orf:invalidArgument", 'Factorization must be "li" or "laurent"')
        if Symmetry not in ("auto", "none"):
            raise OrfError("orf:invalidArgument", 'Symmetry must be "auto" or "none"')
        lam = np.atleast_1d(np.asarray(lambda0, dtype=float))
        self.multi_lambda = np.ndim(lambda0) > 0
        self.lambdas = lam
        pols, self.multi_pol = _pol_kinds(pol)
        self.npol = len(pols)
        nL = len(devs[0]["layers"])
        for d in devs:
            if len(d["layers"]) != nL:
                raise OrfError("orf:invalidArgument", "structures must share the layer count")
        # materials
        mats, index = [], {}

        def mid(m):
            k = ("const", m["n"]) if m["kind"] == "const" else ("func", id(m))
            if k not in index:
                index[k] = len(mats)
                mats.append(m)
            return index[k]
        inc = mid(devs[0]["incident"])
        sub = mid(devs[0]["substrate"])
        shapes, verts = [], []
        thick = np.zeros((nL, len(devs)))
        bg = np.zeros((nL, len(devs)), np.int32)
        s0 = np.zeros((nL, len(devs)), np.int32)
        sc = np.zeros((nL, len(devs)), np.int32)
        for b, d in enumerate(devs):
            for j, L in enumerate(d["layers"]):
                thick[j, b] = L["thickness"]
                bg[j, b] = mid(L["background"])
                s0[j, b] = len(shapes)
                sc[j, b] = len(L["shapes"])
                for s in L["shapes"]:
                    t = _SHAPE_TYPES.get(s["type"])
                    if t is None:
                        raise OrfError("orf:invalidArgument", f"unsupported shape {s['type']}")
                    rec = _lib.orf_shape(t, mid(s["mat"]), s["center"][0], s["center"][1], 0.0, 0.0, 0, 0)
                    if t == 0:
                        rec.a = s["radius"]
                    elif t == 1:
                        rec.a, rec.b = s["size"]
                    else:
                        v = np.asarray(s["vertices"], float)
                        rec.vert0 = len(verts)
                        rec.nvert = v.shape[0]
                        verts.extend(v.tolist())
                    shapes.append(rec)
        eps = np.zeros((len(mats), lam.size), complex)
        for i, m in enumerate(mats):
            for l, lm in enumerate(lam):
                eps[i, l] = material_index(m, lm) ** 2
        if isinstance(Nsample, (int, float)) and Nsample == 0:
            Nsx, Nsy = default_sample_counts(devs, Nx, Ny)
        else:
            Nsx, Nsy = int(Nsample[0]), int(Nsample[1])
        if not math.isfinite(Ly):
            Nsy = 1
        # keep buffers alive
        self._lam = np.ascontiguousarray(lam)
        self._polkind = np.array([p[0] for p in pols], np.int32)
        self._jones = np.array([[p[1].real, p[1].imag, p[2].real, p[2].imag] for p in pols], float).ravel()
        self._eps = np.ascontiguousarray(eps.T).reshape(-1)  # material fastest, then lambda
        self._epsd = self._eps.view(float)
        self._thick = np.ascontiguousarray(thick.reshape(-1, order="F"))
        self._bg = np.ascontiguousarray(bg.reshape(-1, order="F"))
        self._s0 = np.ascontiguousarray(s0.reshape(-1, order="F"))
        self._sc = np.ascontiguousarray(sc.reshape(-1, order="F"))
        self._shapes = (_lib.orf_shape * max(1, len(shapes)))(*shapes)
        self._verts = np.ascontiguousarray(np.asarray(verts if verts else [[0.0, 0.0]], float).reshape(-1))
        P = C.POINTER
        pr = _lib.orf_problem()
        pr.Lx, pr.Ly = Lx, Ly
        pr.Nx, pr.Ny = int(Nx), int(Ny)
        pr.Nsx, pr.Nsy = int(Nsx), int(Nsy)
        pr.nLambda = lam.size
        pr.lambda0 = self._lam.ctypes.data_as(P(C.c_double))
        pr.theta, pr.phi = float(theta), float(phi)
        pr.nPol = len(pols)
        pr.polKind = self._polkind.ctypes.data_as(P(C.c_int32))
        pr.jones = self._jones.ctypes.data_as(P(C.c_double))
        pr.nMaterials = len(mats)
        pr.matEps = self._epsd.ctypes.data_as(P(C.c_double))
        pr.incMaterial, pr.subMaterial = inc, sub
        pr.nStruct, pr.nLayers = len(devs), nL
        pr.thickness = self._thick.ctypes.data_as(P(C.c_double))
        pr.background = self._bg.ctypes.data_as(P(C.c_int32))
        pr.shapeStart = self._s0.ctypes.data_as(P(C.c_int32))
        pr.shapeCount = self._sc.ctypes.data_as(P(C.c_int32))
        pr.nShapes = len(shapes)
        pr.shapes = C.cast(self._shapes, P(_lib.orf_shape))
        pr.nVertices = len(verts)
        pr.vertices = self._verts.ctypes.data_as(P(C.c_double))
        pr.factorization = 0 if Factorization == "li" else 1
        pr.symmetry = 0 if Symmetry == "auto" else 1
        pr.gapEps = float(GapEps)
        if Precision not in ("fp64", "mixed", "split"):
            raise OrfError("orf:invalidArgument", 'Precision must be "fp64", "mixed" or "split"')
        pr.precision = {"fp64": 0, "mixed": 1, "split": 2}[Precision]
        pr.batchSize = int(BatchSize)
        if self.pages is not None:
            if self.pages.shape[0] == 0:
                raise OrfError("orf:invalidArgument", "Pages is empty")
            if (self.pages[:, 0].min() < 0 or self.pages[:, 0].max() >= len(devs)
                    or self.pages[:, 1].min() < 0 or self.pages[:, 1].max() >= lam.size):
                raise OrfError("orf:invalidArgument", "Pages index out of range")
            self._period = np.ascontiguousarray(periods.reshape(-1))
            self._pstruct = np.ascontiguousarray(self.pages[:, 0])
            self._plam = np.ascontiguousarray(self.pages[:, 1])
            pr.structPeriod = self._period.ctypes.data_as(P(C.c_double))
            pr.nPages = int(self.pages.shape[0])
            pr.pageStruct = self._pstruct.ctypes.data_as(P(C.c_int32))
            pr.pageLambda = self._plam.ctypes.data_as(P(C.c_int32))
        self.struct = pr
        self.M = (2 * int(Nx) + 1) * (2 * int(Ny) + 1)
        self.B = len(devs)
        self.Nx, self.Ny = int(Nx), int(Ny)
        self.Nsample = (int(Nsx), int(Nsy))


def estimate_memory(problem: Problem, reduced: bool) -> int:
    """Upper bound of the device bytes a plan of ``problem`` will hold."""
    b = C.c_int64()
    _raise(_lib.load_library().orf_plan_estimate(C.byref(problem.struct), 1 if reduced else 0, C.byref(b)))
    return b.value


def _raise(rc: int, pending_exc=None):
    if rc == _lib.ORF_OK:
        return
    if pending_exc is not None:
        raise pending_exc
    raise OrfError(_lib.ERROR_IDS.get(rc, "orf:error"), _lib.last_error())


def _make_progress(cb):
    state = {"exc": None}
    if cb is None:
        return _lib.PROGRESS_FN(0), state

    def f(frac, msg, user):
        try:
            cb(float(frac), msg.decode() if msg else "")
            return 0
        except BaseException as e:  # the callback cancels by raising (rcwa2D.m:28-29)
            state["exc"] = e
            return 1
    return _lib.PROGRESS_FN(f), state


class Plan:
    """A device-resident solve (orf_plan_*): upload once, execute many times."""

    def __init__(self, problem: Problem):
        lib = _lib.load_library()
        self.problem = problem
        self._lib = lib
        h = C.c_void_p()
        rc = lib.orf_plan_create(_lib.context(), C.byref(problem.struct), C.byref(h))
        _raise(rc)
        self.handle = h
        ns, sym, npg = C.c_int32(), C.c_int32(), C.c_int32()
        lib.orf_plan_info(h, C.byref(ns), C.byref(sym), C.byref(npg))
        self.solved_size, self.solve_pages = ns.value, npg.value
        self.symmetric = sym.value == 1                      # mirror-parity sectors
        self.basis = ("full", "mirror", "te-tm")[sym.value]  # te-tm: y-invariant grating, k_y = 0

    def memory(self) -> int:
        """Device bytes this plan holds (tables, results, workspace)."""
        b = C.c_int64()
        _raise(self._lib.orf_plan_memory(self.handle, C.byref(b)))
        return b.value

    def execute(self, Progress=None):
        fn, state = _make_progress(Progress)
        rc = self._lib.orf_plan_execute(self.handle, fn, None)
        _raise(rc, state["exc"])

    def enable_graphs(self, on=True):
        """Capture each chunk's stage chain into a HIP graph and replay it on later executes."""
        self._lib.orf_plan_enable_graphs(self.handle, 1 if on else 0)

    def enable_timing(self, on=True):
        self._lib.orf_plan_enable_timing(self.handle, 1 if on else 0)

    def stage_times(self):
        n = 64
        names = (C.c_char_p * n)()
        ms = (C.c_double * n)()
        cnt = (C.c_int32 * n)()
        k = C.c_int32()
        self._lib.orf_plan_stage_times(self.handle, n, names, ms, cnt, C.byref(k))
        return {names[i].decode(): (ms[i], cnt[i]) for i in range(k.value)}

    def fetch(self, amplitudes=True):
        pr = self.problem
        M, B, L, Pn = pr.M, pr.B, pr.lambdas.size, pr.npol
        if pr.pages is not None:  # one "structure" axis of nPages, no wavelength axis
            B, L = pr.pages.shape[0], 1
        m = np.zeros(M, np.int32)
        n = np.zeros(M, np.int32)
        R = np.zeros((Pn, L, B, M))
        T = np.zeros((Pn, L, B, M))
        Rt = np.zeros((Pn, L, B))
        Tt = np.zeros((Pn, L, B))
        r = np.zeros((Pn, L, B, M, 3), complex) if amplitudes else None
        t = np.zeros((Pn, L, B, M, 3), complex) if amplitudes else None
        P = C.POINTER
        res = _lib.orf_result(m.ctypes.data_as(P(C.c_int32)), n.ctypes.data_as(P(C.c_int32)),
                              _lib.dptr(R), _lib.dptr(T), _lib.dptr(Rt), _lib.dptr(Tt),
                              _lib.dptr(r) if amplitudes else None, _lib.dptr(t) if amplitudes else None)
        _raise(self._lib.orf_plan_fetch(self.handle, C.byref(res)))
        return _pack(pr, m, n, R, T, Rt, Tt, r, t, self.symmetric, self.basis)

    def close(self):
        if getattr(self, "handle", None):
            self._lib.orf_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _pack(pr, m, n, R, T, Rt, Tt, r, t, symmetric, basis="full"):
    """C-order (pol, lambda, B, M[, 3]) -> MATLAB order (M, B[, lambda][, pol])."""
    def fix(a, amp=False):
        if a is None:
            return None
        a = np.moveaxis(a, [0, 1], [-1, -2])  # (B, M[,3], L, P)
        if amp:
            a = np.moveaxis(a, 2, 0)          # (3, B, M, L, P)
            a = np.swapaxes(a, 1, 2)          # (3, M, B, L, P)
        else:
            a = np.swapaxes(a, 0, 1) if a.ndim == 4 else a  # (M, B, L, P)
        if not pr.multi_lambda or pr.pages is not None:
            a = a[..., 0, :]
        if not pr.multi_pol:
            a = a[..., 0]
        return a
    out = {"m": m.astype(int), "n": n.astype(int), "R": fix(R), "T": fix(T), "Rtot": fix(Rt),
           "Ttot": fix(Tt), "r": fix(r, True), "t": fix(t, True), "symmetry": symmetric, "basis": basis,
           "Nsample": pr.Nsample}
    return out


def rcwa2D(devs, lambda0, theta, phi, pol, Nx=5, Ny=None, Nsample=0, Factorization="li",
           GapEps=float("nan"), UseGPU=True, GPUEig=True, BatchSize=0, Symmetry="auto",
           Progress: Optional[Callable] = None, Precision="fp64", Pages=None):
    """orf.rcwa2D (rcwa2D.m:2-35) on the GPU.

    ``UseGPU`` / ``GPUEig`` are accepted for signature compatibility; every
    stage, eig included, runs on the gfx950 device.  ``Pages`` (build
    extension) is a list of (structure, wavelength-index) pairs to solve
    instead of every pair; the structures may then have different periods.
    """
    prob = Problem(devs, lambda0, theta, phi, pol, Nx=Nx, Ny=Ny, Nsample=Nsample,
                   Factorization=Factorization, GapEps=GapEps, Symmetry=Symmetry,
                   BatchSize=BatchSize, Precision=Precision, Pages=Pages)
    plan = Plan(prob)
    try:
        plan.execute(Progress)
        return plan.fetch()
    finally:
        plan.close()


def rcwaFields(dev, lambda0, theta, phi, pol, x, y, z, Nx=5, Ny=None, Nsample=0, Factorization="li",
               GapEps=float("nan")):
    """orf.rcwaFields (rcwaFields.m:1-21) on the GPU.

    Returns E, H (H~ = i eta0 H) as nx x ny x nz x 3 complex arrays, eps at
    the points, the grid, zInterfaces, flux (normalised to the incident flux)
    and ``out`` (the rcwa2D-style result).  z = 0 is the top of the first layer.
    """
    if isinstance(pol, (list, tuple)) and pol and isinstance(pol[0], str):
        raise OrfError("rcwa:pol", "rcwaFields takes one polarisation")
    x = np.ascontiguousarray(np.atleast_1d(np.asarray(x, float)))
    y = np.ascontiguousarray(np.atleast_1d(np.asarray(y, float)))
    z = np.ascontiguousarray(np.atleast_1d(np.asarray(z, float)))
    prob = Problem([dev], float(np.asarray(lambda0).ravel()[0]), theta, phi, pol, Nx=Nx, Ny=Ny, Nsample=Nsample,
                   Factorization=Factorization, GapEps=GapEps, Symmetry="none")
    nx, ny, nz = x.size, y.size, z.size
    E = np.zeros((3, nz, ny, nx), complex)
    H = np.zeros((3, nz, ny, nx), complex)
    flux = np.zeros(nz)
    M = prob.M
    m = np.zeros(M, np.int32)
    n = np.zeros(M, np.int32)
    R = np.zeros((1, 1, 1, M))
    T = np.zeros((1, 1, 1, M))
    Rt = np.zeros((1, 1, 1))
    Tt = np.zeros((1, 1, 1))
    r = np.zeros((1, 1, 1, M, 3), complex)
    t = np.zeros((1, 1, 1, M, 3), complex)
    Pp = C.POINTER
    res = _lib.orf_result(m.ctypes.data_as(Pp(C.c_int32)), n.ctypes.data_as(Pp(C.c_int32)), _lib.dptr(R),
                          _lib.dptr(T), _lib.dptr(Rt), _lib.dptr(Tt), _lib.dptr(r), _lib.dptr(t))
    lib = _lib.load_library()
    rc = lib.orf_fields(_lib.context(), C.byref(prob.struct), nx, _lib.dptr(x), ny, _lib.dptr(y), nz, _lib.dptr(z),
                        _lib.dptr(E), _lib.dptr(H), _lib.dptr(flux), C.byref(res))
    _raise(rc)
    out = _pack(prob, m, n, R, T, Rt, Tt, r, t, False)
    E = np.transpose(E, (3, 2, 1, 0))   # column-major nx x ny x nz x 3 -> numpy axes (x, y, z, c)
    H = np.transpose(H, (3, 2, 1, 0))
    Lx, Ly = dev["period"]
    ztop = np.concatenate([[0.0], np.cumsum([L["thickness"] for L in dev["layers"]])])
    lam = float(np.asarray(lambda0).ravel()[0])
    eps = np.zeros((nx, ny, nz), complex)
    for k, zz in enumerate(z):
        if zz < 0:
            eps[:, :, k] = material_index(dev["incident"], lam) ** 2
        elif zz >= ztop[-1]:
            eps[:, :, k] = material_index(dev["substrate"], lam) ** 2
        else:
            j = int(np.searchsorted(ztop, zz, side="right") - 1)
            eps[:, :, k] = layerEps(dev["layers"][j], [Lx, Ly], x, y, lam)
    return {"E": E, "H": H, "eps": eps, "x": x, "y": y, "z": z, "zInterfaces": ztop, "flux": flux, "out": out}


def rcwa1D(dev, lambda0, theta, phi, pol, N=10, Nsample=0, Factorization="li", GapEps=float("nan")):
    """rcwa2D with Ny = 0 (rcwa1D.m:20-27)."""
    if not (math.isfinite(dev["period"][0]) and not math.isfinite(dev["period"][1])):
        raise OrfError("rcwa1D:period", "rcwa1D needs period [Lx Inf].")
    Ns = Nsample
    if Ns == 0:
        Ns = alignedSampleCount([dev], dev["period"], 1, max(1024, 8 * (4 * N + 1)))
    return rcwa2D(dev, lambda0, theta, phi, pol, Nx=N, Nsample=(Ns, 1), Factorization=Factorization,
                  GapEps=GapEps)


def buildPhaseLibrary(makeCell: Callable, params: Sequence, lambda0, N=10, Pol="s", CheckN=0,
                      UseGPU=True, Progress=None, Symmetry="auto"):
    """Zero-order co-polarised transmission of a unit-cell sweep (buildPhaseLibrary.m:2-17).

    One batched rcwa2D at normal incidence; phase unwrapped relative to the
    first cell; ``CheckN > 0`` reruns at that truncation and stores the
    per-cell phase difference in ``meta``.
    """
    devs = [makeCell(p) for p in params]

    def run(NN, prog):
        o = rcwa2D(devs, lambda0, 0.0, 0.0, Pol, Nx=NN, Progress=prog, Symmetry=Symmetry)
        k0 = int(np.flatnonzero((o["m"] == 0) & (o["n"] == 0))[0])
        comp = 1 if Pol == "s" else 0
        return o["t"][comp, k0, :], o["T"][k0, :]
    t, T = run(N, Progress)
    phase = np.unwrap(np.angle(t / t[0]))
    lib = {"param": np.asarray(params), "t": t, "T": T, "phase": phase,
           "span": float(phase.max() - phase.min()),
           "meta": {"lambda0": lambda0, "N": N, "pol": Pol, "period": devs[0]["period"]}}
    if CheckN > 0:
        t2, T2 = run(CheckN, None)
        dphi = np.abs(np.angle(t2 / t))
        lib["meta"].update({"CheckN": CheckN, "phaseDiff": dphi, "maxPhaseDiff": float(dphi.max()),
                            "maxTDiff": float(np.max(np.abs(T2 - T)))})
    return lib


def convergeN(fun: Callable, Tol=1e-3, Ns=range(3, 26, 2), Progress=None):
    """Raise N until two successive results differ by < Tol (convergeN.m:20-44).

    Returns ``N`` (the first N that met Tol, or the last tried), ``value``
    (fun(N)), ``converged`` and ``history``, the MATLAB table's columns
    (N, change, seconds) as lists (convergeN.m:39-40).  An unconverged ladder
    warns with identifier ``orf:convergeN:notConverged`` (convergeN.m:41-44).
    """
    import time
    import warnings
    if not Tol > 0:
        raise OrfError("orf:invalidArgument", "Tol must be positive")
    Ns = [int(v) for v in Ns]
    if not Ns or min(Ns) < 0:
        raise OrfError("orf:invalidArgument", "Ns must be a non-empty list of non-negative integers")
    prev = None
    hist = {"N": [], "change": [], "seconds": []}
    res = {"converged": False}
    for k, N in enumerate(Ns):
        if Progress is not None:
            Progress(k / len(Ns), f"N = {N}")
        t0 = time.perf_counter()
        q = np.atleast_1d(np.asarray(fun(N)))
        secs = time.perf_counter() - t0
        change = float(np.max(np.abs(q.ravel() - prev.ravel()))) if prev is not None else float("nan")
        hist["N"].append(N)
        hist["change"].append(change)
        hist["seconds"].append(secs)
        prev = q
        res["N"], res["value"] = N, q
        if change < Tol:
            res["converged"] = True
            break
    res["history"] = hist
    if not res["converged"]:
        warnings.warn(OrfWarning("orf:convergeN:notConverged",
                                 f"Not converged to {Tol:g} by N = {res['N']} "
                                 f"(last change {hist['change'][-1]:.2g})."))
    return res
