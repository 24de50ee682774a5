"""Device descriptions for the 2D RCWA unit-cell solve.

Plain-dict mirrors of the reference's model layer, with the same names,
argument order and field meaning, so code written against the MATLAB toolbox
reads the same here:

* ``materialConst``  -> toolbox/+orf/materialConst.m:1-12
* ``makeStructure``  -> toolbox/+orf/makeStructure.m:1-20
* ``makeLayer``      -> toolbox/+orf/makeLayer.m:1-14
* ``shapeCircle``    -> toolbox/+orf/shapeCircle.m:1-13
* ``shapeRect``      -> toolbox/+orf/shapeRect.m:1-16
* ``shapePolygon``   -> toolbox/+orf/shapePolygon.m:1-18
* ``slantedLayers``  -> toolbox/+orf/slantedLayers.m:1-62
* ``cellGrid``       -> toolbox/+orf/cellGrid.m:1-20
* ``alignedSampleCount`` -> toolbox/+orf/alignedSampleCount.m:1-42

Shapes carry their parameters instead of a function handle; the inside
predicate of each type is fixed (strict inequalities for circle and
rectangle, edge-inclusive for polygons, as MATLAB's ``inpolygon``).
The permittivity grid itself is painted on the GPU from these descriptors.
"""
from __future__ import annotations

import math
from typing import Callable, Sequence, Union

import numpy as np

Number = Union[float, complex]


class OrfError(ValueError):
    """Error carrying the reference's MATLAB error identifier (e.g. ``rcwa:pol``)."""

    def __init__(self, ident: str, msg: str):
        super().__init__(f"{ident}: {msg}")
        self.identifier = ident


class OrfWarning(UserWarning):
    """Warning carrying the reference's MATLAB warning identifier
    (e.g. ``orf:convergeN:notConverged``)."""

    def __init__(self, ident: str, msg: str):
        super().__init__(f"{ident}: {msg}")
        self.identifier = ident


# ---------------------------------------------------------------- materials
def materialConst(n: Number, name: str = "const") -> dict:
    """Non-dispersive material with constant complex index (materialConst.m:9-11)."""
    return {"name": str(name), "kind": "const", "n": complex(n)}


def material(nk: Callable[[float], Number], name: str = "dispersive") -> dict:
    """Material whose index is given by a callable n(lambda0 [um]).

    The callable is evaluated on the host, once per wavelength, to one scalar;
    the device only ever sees eps = n^2 per (material, wavelength).
    """
    return {"name": str(name), "kind": "func", "nk": nk}


def material_index(mat: dict, lambda0: float) -> complex:
    if mat["kind"] == "const":
        return mat["n"]
    return complex(mat["nk"](float(lambda0)))


# ---------------------------------------------------------------- structure
def makeStructure(period: Sequence[float], incident: dict, substrate: dict,
                  layers: Sequence[dict] = ()) -> dict:
    """Solver-independent device description (makeStructure.m:16-19).

    Layers are listed from the incident side toward the substrate (+z).
    ``period = [Lx, Ly]``; ``Ly = inf`` describes a y-invariant grating.
    """
    period = [float(period[0]), float(period[1])]
    if not (period[0] > 0 and period[1] > 0):
        raise OrfError("orf:makeStructure:period", "period must be positive")
    return {"period": period, "incident": incident, "substrate": substrate,
            "layers": list(layers)}


def makeLayer(thickness: float, background: dict, shapes: Sequence[dict] = ()) -> dict:
    """One z-uniform layer (makeLayer.m:11-13); shapes paint in order."""
    if thickness < 0:
        raise OrfError("orf:makeLayer:thickness", "thickness must be non-negative")
    return {"thickness": float(thickness), "background": background,
            "shapes": list(shapes)}


def shapeCircle(r: float, mat: dict, center: Sequence[float] = (0.0, 0.0)) -> dict:
    """Circle, inside test dx^2 + dy^2 < r^2 (shapeCircle.m:12)."""
    if not r > 0:
        raise OrfError("orf:shapeCircle:radius", "radius must be positive")
    return {"type": "circle", "mat": mat, "center": [float(center[0]), float(center[1])],
            "radius": float(r)}


def shapeRect(wx: float, wy: float, mat: dict, center: Sequence[float] = (0.0, 0.0)) -> dict:
    """Axis-aligned rectangle, inside |dx| < wx/2 & |dy| < wy/2 (shapeRect.m:15)."""
    if not (wx > 0 and wy > 0):
        raise OrfError("orf:shapeRect:size", "rectangle sides must be positive")
    return {"type": "rect", "mat": mat, "center": [float(center[0]), float(center[1])],
            "size": [float(wx), float(wy)]}


def shapePolygon(xy, mat: dict, center: Sequence[float] = (0.0, 0.0)) -> dict:
    """Polygon relative to ``center``; edge-inclusive inside test (shapePolygon.m:17)."""
    xy = np.asarray(xy, dtype=float).reshape(-1, 2)
    if xy.shape[0] < 3:
        raise OrfError("orf:shapePolygon:vertices", "A polygon needs at least 3 vertices.")
    return {"type": "polygon", "mat": mat, "center": [float(center[0]), float(center[1])],
            "vertices": xy.copy()}


def shapeCross(armWidth: float, armLength: float, mat: dict,
               center: Sequence[float] = (0.0, 0.0)) -> list:
    """A cross as two rectangles painted in order (the reference has no cross type)."""
    return [shapeRect(armLength, armWidth, mat, center),
            shapeRect(armWidth, armLength, mat, center)]


# ---------------------------------------------------------------- slanted walls
def _signed_area(xy: np.ndarray) -> float:
    return 0.5 * float(np.sum(xy[:, 0] * np.roll(xy[:, 1], -1) - np.roll(xy[:, 0], -1) * xy[:, 1]))


def _segments_meet(p1, p2, q1, q2) -> bool:
    """Closed segments p1-p2 and q1-q2 share a point (touching and collinear overlap included)."""
    def orient(a, b, c):
        return (b[0] - a[0]) * (c[1] - a[1]) - (b[1] - a[1]) * (c[0] - a[0])

    def on(a, b, c):  # c on segment a-b, given collinear
        return min(a[0], b[0]) <= c[0] <= max(a[0], b[0]) and min(a[1], b[1]) <= c[1] <= max(a[1], b[1])
    d1, d2 = orient(q1, q2, p1), orient(q1, q2, p2)
    d3, d4 = orient(p1, p2, q1), orient(p1, p2, q2)
    if d1 * d2 < 0 and d3 * d4 < 0:
        return True
    return ((d1 == 0 and on(q1, q2, p1)) or (d2 == 0 and on(q1, q2, p2)) or
            (d3 == 0 and on(p1, p2, q1)) or (d4 == 0 and on(p1, p2, q2)))


def _is_simple(xy: np.ndarray) -> bool:
    k = xy.shape[0]
    for i in range(k):
        for j in range(i + 2, k):
            if i == 0 and j == k - 1:
                continue  # neighbours through the closing edge
            if _segments_meet(xy[i], xy[(i + 1) % k], xy[j], xy[(j + 1) % k]):
                return False
    return True


def _polygon_offset_inward(xy: np.ndarray, d: float):
    """Move every edge of a simple polygon inward by d with miter joins.

    Edges whose offset runs backwards (they shrink to nothing before d) are
    dropped and the neighbours re-joined, so a convex polygon gives exactly
    the intersection of the inward-shifted half-planes.  Returns the vertices,
    None when nothing is left, or raises orf:slantedLayers:split when the
    offset outline crosses itself, i.e. the region pinches into pieces
    (slantedLayers.m:54-57).
    """
    xy = np.asarray(xy, float)
    if _signed_area(xy) < 0:         # make counter-clockwise
        xy = xy[::-1]
    edges = []                       # (point on the offset line, unit direction)
    k = xy.shape[0]
    for i in range(k):
        p, q = xy[i], xy[(i + 1) % k]
        t = q - p
        L = math.hypot(t[0], t[1])
        if L == 0.0:
            continue
        t = t / L
        edges.append((p + d * np.array([-t[1], t[0]]), t))   # inward normal for CCW
    while len(edges) >= 3:
        k = len(edges)
        out = np.empty((k, 2))
        for i in range(k):           # vertex i joins edge i-1 and edge i
            (p1, t1), (p2, t2) = edges[i - 1], edges[i]
            den = t1[0] * t2[1] - t1[1] * t2[0]
            if abs(den) < 1e-15:
                out[i] = p2
                continue
            s_ = ((p2[0] - p1[0]) * t2[1] - (p2[1] - p1[1]) * t2[0]) / den
            out[i] = p1 + s_ * t1
        inverted = [i for i in range(k) if float(np.dot(out[(i + 1) % k] - out[i], edges[i][1])) <= 0.0]
        if not inverted:
            if _signed_area(out) <= 0.0:
                return None
            if not _is_simple(out):
                raise OrfError("orf:slantedLayers:split", "The polygon splits into pieces at this sidewall angle.")
            return out
        # drop the edge that vanished first: the one whose offset is most reversed
        worst = min(inverted, key=lambda i: float(np.dot(out[(i + 1) % k] - out[i], edges[i][1])))
        edges.pop(worst)
    return None


def _shrink(shape: dict, d: float) -> list:
    t = shape["type"]
    if t == "circle":
        r = shape["radius"] - d
        return [shapeCircle(r, shape["mat"], shape["center"])] if r > 0 else []
    if t == "rect":
        sx, sy = shape["size"][0] - 2 * d, shape["size"][1] - 2 * d
        return [shapeRect(sx, sy, shape["mat"], shape["center"])] if (sx > 0 and sy > 0) else []
    if t == "polygon":
        v = _polygon_offset_inward(shape["vertices"], d)
        return [] if v is None else [shapePolygon(v, shape["mat"], shape["center"])]
    raise OrfError("orf:slantedLayers:shape", f'Unsupported shape "{t}".')


def slantedLayers(thickness: float, background: dict, shape, sidewallDeg: float,
                  NSlices: int = 8, Base: str = "bottom") -> list:
    """Staircase of NSlices layers for a sloped sidewall (slantedLayers.m:26-40).

    Every edge moves inward by h' cot(sidewallDeg) at the slice mid-height h'
    above the base; ``shape`` may be one shape or a list (e.g. a cross).
    """
    if not (1 <= sidewallDeg <= 179):
        raise OrfError("orf:slantedLayers:angle", "sidewallDeg must be in [1, 179]")
    if Base not in ("bottom", "top"):
        raise OrfError("orf:slantedLayers:base", 'Base must be "bottom" or "top"')
    shapes = shape if isinstance(shape, (list, tuple)) else [shape]
    N = int(NSlices)
    dh = thickness / N
    layers = []
    cot = math.cos(math.radians(sidewallDeg)) / math.sin(math.radians(sidewallDeg))
    for k in range(1, N + 1):
        z_mid = (k - 0.5) * dh
        above = thickness - z_mid if Base == "bottom" else z_mid
        inset = above * cot
        sl = []
        for s in shapes:
            sl.extend(_shrink(s, inset))
        layers.append(makeLayer(dh, background, sl))
    return layers


# ---------------------------------------------------------------- sampling
def cellGrid(period: Sequence[float], Nx: int, Ny: int = 1):
    """Cell-centred sample points (cellGrid.m:14-20): x_i = -L/2 + L/N (i + 1/2)."""
    def axis(L, N):
        if not math.isfinite(L):
            return np.zeros(1)
        return (-L / 2) + (L / N) * (np.arange(N) + 0.5)
    return axis(period[0], Nx), axis(period[1], Ny)


def alignedSampleCount(devs: Sequence[dict], period: Sequence[float], dim: int, NsMin: int) -> int:
    """Smallest Ns in [NsMin, 4 NsMin] putting every rectangle edge on a pixel
    boundary (alignedSampleCount.m:16-41); ``dim`` is 1 (x) or 2 (y)."""
    L = period[dim - 1]
    edges = []
    for dev in devs:
        for layer in dev["layers"]:
            for s in layer["shapes"]:
                if s["type"] == "rect" and s["size"][dim - 1] < L:
                    c = s["center"][dim - 1]
                    h = s["size"][dim - 1] / 2
                    edges += [c - h, c + h]
    if not edges or not math.isfinite(L):
        return int(NsMin)
    u = np.mod(np.asarray(edges) / L + 0.5, 1.0)
    u = np.unique(np.round(u / 1e-12)) * 1e-12
    for cand in range(int(NsMin), 4 * int(NsMin) + 1):
        f = u * cand
        if np.all(np.abs(f - np.round(f)) < 1e-6):
            return cand
    return int(NsMin)


def default_sample_counts(devs: Sequence[dict], Nx: int, Ny: int) -> tuple:
    """Automatic [Nsx Nsy] for ``Nsample = 0``.

    The reference's own automatic rule is not published; this package uses
    NsMin = max(256, 8 (4N + 1)) per periodic axis, edge-aligned for
    rectangles, and a single sample along an infinite period.
    """
    period = devs[0]["period"]
    shared = all(list(d["period"]) == list(period) for d in devs)
    out = []
    for dim, N in ((1, Nx), (2, Ny)):
        ns_min = max(256, 8 * (4 * N + 1))
        if not math.isfinite(period[dim - 1]):
            out.append(1)
        elif shared:
            out.append(alignedSampleCount(devs, period, dim, ns_min))
        else:  # structures of different periods cannot all be edge-aligned by one count
            out.append(ns_min)
    return tuple(out)


# ---------------------------------------------------------------- permittivity at points
def _wrap(d, L):
    """Displacement mapped into [-L/2, L/2) (periodic unit cell, layerEps.m:40-45)."""
    if not math.isfinite(L):
        return d
    t = d + L / 2
    return t - np.floor(t / L) * L - L / 2


def _in_polygon(px, py, v):
    """Edge-inclusive even-odd test of points against polygon vertices v (K x 2)."""
    inside = np.zeros(px.shape, bool)
    on = np.zeros(px.shape, bool)
    k = v.shape[0]
    for i in range(k):
        x1, y1 = v[i]
        x2, y2 = v[(i + 1) % k]
        cross = (x2 - x1) * (py - y1) - (y2 - y1) * (px - x1)
        box = (np.minimum(x1, x2) <= px) & (px <= np.maximum(x1, x2)) & \
              (np.minimum(y1, y2) <= py) & (py <= np.maximum(y1, y2))
        on |= (np.abs(cross) <= 1e-12 * max(1.0, abs(x2 - x1) + abs(y2 - y1))) & box
        crosses = (y1 > py) != (y2 > py)
        with np.errstate(divide="ignore", invalid="ignore"):
            xs = x1 + (py - y1) * (x2 - x1) / (y2 - y1)
        inside ^= crosses & (px < xs)
    return inside | on


def layerEps(layer: dict, period: Sequence[float], x, y, lambda0: float) -> np.ndarray:
    """eps = n^2 of a layer at the points x (nx) by y (ny): background, then the
    shapes painted in order with periodic wrap (layerEps.m:19-45)."""
    X, Y = np.meshgrid(np.atleast_1d(np.asarray(x, float)), np.atleast_1d(np.asarray(y, float)), indexing="ij")
    eps = np.full(X.shape, material_index(layer["background"], lambda0) ** 2, complex)
    for sh in layer["shapes"]:
        dx = _wrap(X - sh["center"][0], period[0])
        dy = _wrap(Y - sh["center"][1], period[1])
        if sh["type"] == "circle":
            mask = dx * dx + dy * dy < sh["radius"] ** 2
        elif sh["type"] == "rect":
            mask = (np.abs(dx) < sh["size"][0] / 2) & (np.abs(dy) < sh["size"][1] / 2)
        else:
            mask = _in_polygon(dx, dy, np.asarray(sh["vertices"], float))
        eps[mask] = material_index(sh["mat"], lambda0) ** 2
    return eps
