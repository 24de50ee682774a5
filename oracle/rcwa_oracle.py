"""FP64 CPU restatement of the reference's 2D RCWA unit-cell solve.

TEST INFRASTRUCTURE ONLY.  Nothing in the product path may import this module:
only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg
use it, and only as the reference the results are compared with / the timed CPU
baseline.

The reference (MoonUsagi/OpticRCWAFDTD, MATLAB) ships the solver core as
P-code, which is never read.  This module restates the algorithm from the
open helpers and docstrings it is built on, each cited as file:line relative
to the reference root.  It is pinned (tests/test_oracle_*.py) by

* the RETICOLO V10 golden vectors held in tests/TestReticolo.m:21-78,
* TMM equality and the TMM closed forms (tests/TestRCWA2D.m:25-33,
  tests/TestRCWA1D.m:31-58, tests/TestTMM.m:15-98),
* the invariants of tests/TestRCWA2D.m:49-121 and tests/TestRCWA1D.m:60-138,
* literature values (tests/TestLiteratureRCWA.m:12-28).

Arrays are page-batched with the batch on the LEADING axis (numpy stacking),
where the reference puts pages on dim 3; matrices are otherwise identical.
"""
from __future__ import annotations

import math
from typing import Callable, Optional, Sequence

import numpy as np
import scipy.linalg as sla


class OracleError(ValueError):
    def __init__(self, ident: str, msg: str):
        super().__init__(f"{ident}: {msg}")
        self.identifier = ident


# ------------------------------------------------------------------ helpers
def nk(mat: dict, lambda0: float) -> complex:
    """Index of a material at lambda0 (materialConst.m:10)."""
    if mat["kind"] == "const":
        return complex(mat["n"])
    return complex(mat["nk"](float(lambda0)))


def harmonics2d(Nx: int, Ny: int):
    """Order indices, m fastest (harmonics2D.m:4-6)."""
    mm, nn = np.meshgrid(np.arange(-Nx, Nx + 1), np.arange(-Ny, Ny + 1), indexing="ij")
    return mm.ravel(order="F"), nn.ravel(order="F")


def kz_branch(kz2):
    """Outgoing branch: Im >= 0, Re >= 0 when Im == 0 (kzBranch.m:6-8)."""
    kz = np.sqrt(np.asarray(kz2, dtype=complex))
    flip = (kz.imag < 0) | ((kz.imag == 0) & (kz.real < 0))
    return np.where(flip, -kz, kz)


def kz_nonzero(kz):
    """Nudge exactly grazing orders off zero (kzNonzero.m:8-9)."""
    kz = np.array(kz, dtype=complex, copy=True)
    tiny = 1e-8
    kz[np.abs(kz) < tiny] = tiny * (1 + 1j)
    return kz


def fourier_weights(Ns: int, N: int):
    """(4N+1) x Ns exact staircase Fourier weights (fourierWeights.m:7-12)."""
    h = np.arange(-2 * N, 2 * N + 1, dtype=float)
    xj = (np.arange(Ns) + 0.5) / Ns - 0.5
    u = np.pi * h / Ns
    s = np.ones_like(u)
    nz = u != 0
    s[nz] = np.sin(u[nz]) / u[nz]
    return s[:, None] * np.exp(-2j * np.pi * h[:, None] * xj[None, :]) / Ns


def conv_matrix_1d(f, N: int):
    """Toeplitz C(m,n) = f_{m-n} (convMatrix1D.m:13-18)."""
    f = np.asarray(f, dtype=complex).ravel()
    if f.size < 4 * N + 1:
        raise OracleError("convMatrix1D:sampling", f"Need at least {4 * N + 1} samples, got {f.size}.")
    fh = fourier_weights(f.size, N) @ f
    mid = 2 * N
    return sla.toeplitz(fh[mid:], fh[mid::-1])


def fourier_coeffs_2d(f, Nx: int, Ny: int):
    """f_{h,k}, h in [-2Nx, 2Nx], k in [-2Ny, 2Ny]: separable exact weights."""
    Nsx, Nsy = f.shape
    return fourier_weights(Nsx, Nx) @ f @ fourier_weights(Nsy, Ny).T


def conv_matrix_2d(f, Nx: int, Ny: int):
    """Laurent C(p,q) = f_{m_p - m_q, n_p - n_q} (convMatrix2D.m:2-4)."""
    F2 = fourier_coeffs_2d(np.asarray(f, dtype=complex), Nx, Ny)
    m, n = harmonics2d(Nx, Ny)
    return F2[(m[:, None] - m[None, :]) + 2 * Nx, (n[:, None] - n[None, :]) + 2 * Ny]


def conv_matrices_li_2d(f, Nx: int, Ny: int):
    """Li (1997) Exx, Eyy (convMatricesLi2D.m:2-5).

    Exx: inverse rule along x at every y-sample, then Laurent along y.
    Eyy: inverse rule along y at every x-sample, then Laurent along x.
    """
    f = np.asarray(f, dtype=complex)
    Nsx, Nsy = f.shape
    m, n = harmonics2d(Nx, Ny)
    Fx = fourier_weights(Nsx, Nx)
    Fy = fourier_weights(Nsy, Ny)
    # A[j] = inv(Toeplitz_x(1/eps(:, y_j)))  (2Nx+1)^2 per y-sample
    A = np.stack([np.linalg.inv(conv_matrix_1d(1.0 / f[:, j], Nx)) for j in range(Nsy)])
    # B[i] = inv(Toeplitz_y(1/eps(x_i, :)))  (2Ny+1)^2 per x-sample
    Bm = np.stack([np.linalg.inv(conv_matrix_1d(1.0 / f[i, :], Ny)) for i in range(Nsx)])
    im = m + Nx
    inn = n + Ny
    dn = (n[:, None] - n[None, :]) + 2 * Ny
    dm = (m[:, None] - m[None, :]) + 2 * Nx
    # Exx(p,q) = sum_j Fy(n_p - n_q, j) A_j(m_p, m_q)
    Aq = A[:, im[:, None], im[None, :]]                # Nsy x M x M
    Exx = np.einsum("pqj,jpq->pq", Fy[dn, :], Aq)
    Bq = Bm[:, inn[:, None], inn[None, :]]             # Nsx x M x M
    Eyy = np.einsum("pqi,ipq->pq", Fx[dm, :], Bq)
    return Exx, Eyy


# ------------------------------------------------------------------ modes
def modes_homogeneous(kx, ky, epsr):
    """Analytic modes of a uniform medium (modesHomogeneous.m:6-12)."""
    M = kx.size
    kz = kz_nonzero(kz_branch(epsr - kx ** 2 - ky ** 2))
    lam = -1j * np.concatenate([kz, kz])
    Q = np.block([[np.diag(kx * ky), np.diag(epsr - kx ** 2)],
                  [np.diag(ky ** 2 - epsr), -np.diag(ky * kx)]]).astype(complex)
    W = np.eye(2 * M, dtype=complex)
    V = Q / lam[None, :]
    return W, V, lam, Q


def pq_matrices(Econv, Exx, Eyy, kx, ky):
    """P, Q, EpsInv of a patterned layer (pqMatrices.m:2-4; restated).

    d[s]/dz' = P[u], d[u]/dz' = Q[s], s = [Ex; Ey], u = [H~x; H~y].
    """
    M = kx.size
    I = np.eye(M)
    Ei = np.linalg.inv(Econv)
    KEK = lambda a, b: a[:, None] * Ei * b[None, :]
    P = np.block([[KEK(kx, ky), I - KEK(kx, kx)],
                  [KEK(ky, ky) - I, -KEK(ky, kx)]])
    Q = np.block([[np.diag(kx * ky), Eyy - np.diag(kx ** 2)],
                  [np.diag(ky ** 2) - Exx, -np.diag(ky * kx)]])
    return P, Q, Ei


def modes_layer(P, Q):
    """eig(P*Q) -> W, lam = -i kzBranch(-eig), V = Q W lam^-1 (modesLayer.m:2-4)."""
    mu, W = np.linalg.eig(P @ Q)
    lam = -1j * kz_branch(-mu)
    V = (Q @ W) / lam[..., None, :]
    return W, V, lam


# ------------------------------------------------------------------ S-matrices
def smatrix_layer(W, V, lam, W0, V0, kL):
    """Stable symmetric layer S-matrix (smatrixLayer.m:6-18)."""
    WiW0 = np.linalg.solve(W, np.broadcast_to(W0, W.shape))
    ViV0 = np.linalg.solve(V, np.broadcast_to(V0, V.shape))
    A = WiW0 + ViV0
    B = WiW0 - ViV0
    X = np.exp(-lam * np.asarray(kL)[..., None])          # diagonal as a vector
    XB = X[..., :, None] * B
    n = A.shape[-1]
    Ai = np.linalg.solve(A, np.concatenate([XB, X[..., :, None] * A, B], axis=-1))
    AiXB, AiXA, AiB = Ai[..., :n], Ai[..., n:2 * n], Ai[..., 2 * n:]
    D = A - XB @ AiXB
    S = np.linalg.solve(D, np.concatenate([XB @ AiXA - B, X[..., :, None] * (A - B @ AiB)], axis=-1))
    S11, S12 = S[..., :n], S[..., n:]
    return {"S11": S11, "S12": S12, "S21": S12, "S22": S11}


def _rdiv(Bm, A):
    """B / A = B A^-1 (pagemrdivide)."""
    return np.swapaxes(np.linalg.solve(np.swapaxes(A, -1, -2), np.swapaxes(Bm, -1, -2)), -1, -2)


def smatrix_region(W, V, W0, V0, side: str):
    """Half-space S-matrix (smatrixRegion.m:8-20)."""
    A = np.linalg.solve(W0, W) + np.linalg.solve(V0, V)
    B = np.linalg.solve(W0, W) - np.linalg.solve(V0, V)
    I = np.eye(A.shape[-1])
    AiB = np.linalg.solve(A, B)
    if side == "ref":
        return {"S11": -AiB, "S12": 2 * np.linalg.inv(A), "S21": 0.5 * (A - B @ AiB),
                "S22": _rdiv(B, A)}
    return {"S11": _rdiv(B, A), "S12": 0.5 * (A - B @ AiB), "S21": 2 * np.linalg.inv(A),
            "S22": -AiB}


def smatrix_star(SA, SB):
    """Redheffer star product, pages broadcast (smatrixStar.m:4-10)."""
    n = SA["S11"].shape[-1]
    I = np.eye(n)
    D = _rdiv(np.broadcast_to(SA["S12"], np.broadcast_shapes(SA["S12"].shape, SB["S11"].shape)),
              I - SB["S11"] @ SA["S22"])
    F = _rdiv(np.broadcast_to(SB["S21"], np.broadcast_shapes(SB["S21"].shape, SA["S22"].shape)),
              I - SA["S22"] @ SB["S11"])
    return {"S11": SA["S11"] + D @ SB["S11"] @ SA["S21"],
            "S12": D @ SB["S12"],
            "S21": F @ SA["S21"],
            "S22": SB["S22"] + F @ SA["S22"] @ SB["S12"]}


# ------------------------------------------------------------------ source
def polarization_vector(khat, pol):
    """Unit incident E for "s", "p" or Jones [aTE aTM] (polarizationVector.m:4-20)."""
    if isinstance(pol, str):
        if pol not in ("s", "p"):
            raise OracleError("rcwa:pol", 'pol must be "s", "p" or [aTE aTM].')
        jones = np.array([pol == "s", pol == "p"], dtype=complex)
    else:
        jones = np.asarray(pol, dtype=complex).ravel()
        if jones.size != 2:
            raise OracleError("rcwa:pol", "Jones vector must be [aTE aTM].")
    khat = np.asarray(khat, float)
    if abs(khat[0]) < 1e-12 and abs(khat[1]) < 1e-12:
        aTE = np.array([0.0, 1.0, 0.0])
    else:
        aTE = np.cross([0.0, 0.0, 1.0], khat)
        aTE = aTE / np.linalg.norm(aTE)
    aTM = np.cross(aTE, khat)
    E = jones[0] * aTE + jones[1] * aTM
    return E / np.linalg.norm(E)


# ------------------------------------------------------------------ geometry
def cell_grid(period, Nx: int, Ny: int = 1):
    """Cell-centred samples (cellGrid.m:14-20)."""
    def axis(L, N):
        if not math.isfinite(L):
            return np.zeros(1)
        return (-L / 2) + (L / N) * (np.arange(N) + 0.5)
    return axis(period[0], Nx), axis(period[1], Ny)


def _wrap(d, L):
    """Map a displacement into [-L/2, L/2) (layerEps.m:40-45)."""
    if not math.isfinite(L):
        return d
    t = d + L / 2
    t = t - np.floor(t / L) * L
    return t - L / 2


def _inside(shape, dx, dy):
    t = shape["type"]
    if t == "circle":                      # shapeCircle.m:12
        r = shape["radius"]
        return dx * dx + dy * dy < r * r
    if t == "rect":                        # shapeRect.m:15
        wx, wy = shape["size"]
        return (np.abs(dx) < wx / 2) & (np.abs(dy) < wy / 2)
    raise OracleError("orf:shape", f"unsupported shape {t}")


def _inpolygon(px, py, vx, vy):
    """Edge-inclusive point-in-polygon (crossing number + on-edge test)."""
    px = np.asarray(px, float)
    py = np.asarray(py, float)
    inside = np.zeros(px.shape, bool)
    on = np.zeros(px.shape, bool)
    k = len(vx)
    for i in range(k):
        x1, y1, x2, y2 = vx[i], vy[i], vx[(i + 1) % k], vy[(i + 1) % k]
        cross = (x2 - x1) * (py - y1) - (y2 - y1) * (px - x1)
        within = (np.minimum(x1, x2) <= px) & (px <= np.maximum(x1, x2)) & \
                 (np.minimum(y1, y2) <= py) & (py <= np.maximum(y1, y2))
        on |= (np.abs(cross) <= 1e-12 * max(1.0, abs(x2 - x1) + abs(y2 - y1))) & within
        cond = ((y1 > py) != (y2 > py))
        with np.errstate(divide="ignore", invalid="ignore"):
            xint = x1 + (py - y1) * (x2 - x1) / (y2 - y1)
        inside ^= cond & (px < xint)
    return inside | on


def layer_eps(layer, period, x, y, lambda0):
    """Paint background then shapes in order, periodic wrap (layerEps.m:19-38)."""
    X, Y = np.meshgrid(x, y, indexing="ij")
    e = np.full(X.shape, nk(layer["background"], lambda0) ** 2, dtype=complex)
    for s in layer["shapes"]:
        dx = _wrap(X - s["center"][0], period[0])
        dy = _wrap(Y - s["center"][1], period[1])
        if s["type"] == "polygon":
            v = s["vertices"]
            mask = _inpolygon(dx, dy, v[:, 0], v[:, 1])
        else:
            mask = _inside(s, dx, dy)
        e[mask] = nk(s["mat"], lambda0) ** 2
    return e


def aligned_sample_count(devs, period, dim, NsMin):
    """alignedSampleCount.m:16-41."""
    L = period[dim - 1]
    edges = []
    for d in devs:
        for layer in d["layers"]:
            for s in layer["shapes"]:
                if s["type"] == "rect" and s["size"][dim - 1] < L:
                    edges += [s["center"][dim - 1] - s["size"][dim - 1] / 2,
                              s["center"][dim - 1] + s["size"][dim - 1] / 2]
    if not edges or not math.isfinite(L):
        return int(NsMin)
    u = np.mod(np.asarray(edges) / L + 0.5, 1.0)
    for cand in range(int(NsMin), 4 * int(NsMin) + 1):
        f = u * cand
        if np.all(np.abs(f - np.round(f)) < 1e-6):
            return cand
    return int(NsMin)


def default_sample_counts(devs, Nx, Ny):
    """Automatic Nsample of the GPU library: NsMin = max(256, 8(4N+1)), edge-aligned."""
    period = devs[0]["period"]
    out = []
    for dim, N in ((1, Nx), (2, Ny)):
        if not math.isfinite(period[dim - 1]):
            out.append(1)
        else:
            out.append(aligned_sample_count(devs, period, dim, max(256, 8 * (4 * N + 1))))
    return tuple(out)


# ------------------------------------------------------------------ symmetry
def symmetry_basis(m, n, pol):
    """Orthonormal parity-sector bases UE, UH (2M x Mr) (symmetryBasis.m:2-15).

    pol "p": Ex, Hy even-even; Ey, Hx odd-odd.  pol "s": Ey, Hx even-even;
    Ex, Hy odd-odd.  Columns: even-even sector then odd-odd sector, each
    ordered by (m >= 0, n >= 0) with m fastest.
    """
    M = m.size
    index = {(int(a), int(b)): i for i, (a, b) in enumerate(zip(m, n))}
    Nx, Ny = int(m.max()), int(n.max())
    ee, oo = [], []
    for b in range(0, Ny + 1):
        for a in range(0, Nx + 1):
            v = np.zeros(M)
            orbit = {(a, b), (-a, b), (a, -b), (-a, -b)}
            for (p, q) in orbit:
                v[index[(p, q)]] = 1.0
            ee.append(v / np.linalg.norm(v))
            if a > 0 and b > 0:
                w = np.zeros(M)
                for (p, q) in orbit:
                    w[index[(p, q)]] = np.sign(p) * np.sign(q)
                oo.append(w / np.linalg.norm(w))
    Uee = np.array(ee).T if ee else np.zeros((M, 0))
    Uoo = np.array(oo).T if oo else np.zeros((M, 0))
    Z = lambda a, b: np.zeros((a.shape[0], b.shape[1]))
    if pol == "p":
        UE = np.block([[Uee, Z(Uee, Uoo)], [Z(Uoo, Uee), Uoo]])   # Ex ee | Ey oo
        UH = np.block([[Z(Uoo, Uee), Uoo], [Uee, Z(Uee, Uoo)]])   # Hy ee | Hx oo
    else:
        UE = np.block([[Z(Uee, Uee), Uoo], [Uee, Z(Uee, Uoo)]])   # Ey ee | Ex oo
        UH = np.block([[Uee, Z(Uee, Uoo)], [Z(Uoo, Uee), Uoo]])   # Hx ee | Hy oo
    return UE.astype(complex), UH.astype(complex)


def grid_is_symmetric(e):
    """x -> -x and y -> -y mirror symmetry of a cell-centred sample grid."""
    return bool(np.array_equal(e, e[::-1, :]) and np.array_equal(e, e[:, ::-1]))


# ------------------------------------------------------------------ driver
def rcwa2d(devs, lambda0, theta, phi, pol, Nx=5, Ny=None, Nsample=0, Factorization="li",
           GapEps=float("nan"), Symmetry="auto", BatchSize=0, Progress=None):
    """Batched crossed-grating RCWA (rcwa2D.m:1-35; algorithm restated).

    Returns a dict with m, n (M), R, T (B x M), Rtot, Ttot (B), r, t (B x M x 3).
    """
    if isinstance(devs, dict):
        devs = [devs]
    B = len(devs)
    Lx, Ly = devs[0]["period"]
    if Ny is None:
        Ny = Nx
    if not math.isfinite(Ly):
        Ny = 0
    m, n = harmonics2d(Nx, Ny)
    M = m.size
    lambda0 = float(lambda0)
    k0 = 2 * np.pi / lambda0
    n_inc = nk(devs[0]["incident"], lambda0)
    if abs(n_inc.imag) > 1e-12:
        raise OracleError("rcwa:lossyIncident", "Incident medium must be lossless.")
    n_inc = n_inc.real
    eps_inc = n_inc ** 2
    eps_sub = nk(devs[0]["substrate"], lambda0) ** 2
    st, ct = math.sin(theta), math.cos(theta)
    khat = np.array([st * math.cos(phi), st * math.sin(phi), ct])
    E_inc = polarization_vector(khat, pol)
    kx = n_inc * khat[0] + m * lambda0 / Lx
    ky = n_inc * khat[1] + (n * lambda0 / Ly if math.isfinite(Ly) else 0.0 * n)
    kx = kx.astype(float)
    ky = ky.astype(float)
    if isinstance(Nsample, (int, float)) and Nsample == 0:
        Nsx, Nsy = default_sample_counts(devs, Nx, Ny)
    else:
        Nsx, Nsy = int(Nsample[0]), int(Nsample[1])
    xg, yg = cell_grid([Lx, Ly], Nsx, Nsy)
    nL = len(devs[0]["layers"])
    for d in devs:
        if len(d["layers"]) != nL:
            raise OracleError("rcwa:batch", "structures must share the layer count")

    # grids for patterned layers
    grids = [[(layer_eps(L, [Lx, Ly], xg, yg, lambda0) if L["shapes"] else None)
              for L in d["layers"]] for d in devs]

    use_sym = False
    if Symmetry == "auto" and theta == 0 and isinstance(pol, str) and pol in ("s", "p"):
        use_sym = all(g is None or grid_is_symmetric(g) for gl in grids for g in gl)
    if use_sym:
        UE, UH = symmetry_basis(m, n, pol)
    else:
        UE = UH = None

    def red(Mat, left, right):
        return Mat if left is None else left.conj().T @ Mat @ right

    # gap, reference and transmission media (shared by every page)
    eps_gap = GapEps if not math.isnan(GapEps) else 1.0 + float(np.max(kx ** 2 + ky ** 2))
    _, _, lam0, Q0 = modes_homogeneous(kx, ky, eps_gap)
    def hom(eps):
        W, V, lam, Qh = modes_homogeneous(kx, ky, eps)
        if not use_sym:
            return W, V, lam
        lam_r = np.concatenate([lam[:M] @ UE[:M, :].real ** 2 + lam[M:] @ UE[M:, :].real ** 2])
        Vr = red(Qh, UH, UE) / lam_r[None, :]
        return np.eye(Vr.shape[0], dtype=complex), Vr, lam_r
    W0, V0, _ = hom(eps_gap)
    Wr, Vr_, _ = hom(eps_inc)
    Ws, Vs_, _ = hom(eps_sub)
    S_ref = smatrix_region(Wr, Vr_, W0, V0, "ref")
    S_trn = smatrix_region(Ws, Vs_, W0, V0, "trn")

    bs = B if BatchSize in (0, None) else int(BatchSize)
    chunks = [list(range(i, min(B, i + bs))) for i in range(0, B, bs)]
    S11_all, S21_all = [], []
    nsteps = len(chunks) * max(nL, 1)
    step = 0
    for ch in chunks:
        S = None
        for j in range(nL):
            Ws_, Vs2, lams, kLs = [], [], [], []
            for b in ch:
                L = devs[b]["layers"][j]
                if grids[b][j] is None:
                    W, V, lam = hom(nk(L["background"], lambda0) ** 2)
                else:
                    e = grids[b][j]
                    E = conv_matrix_2d(e, Nx, Ny)
                    if Factorization == "li":
                        Exx, Eyy = conv_matrices_li_2d(e, Nx, Ny)
                    else:
                        Exx = Eyy = E
                    P, Q, _ = pq_matrices(E, Exx, Eyy, kx, ky)
                    W, V, lam = modes_layer(red(P, UE, UH), red(Q, UH, UE))
                Ws_.append(W); Vs2.append(V); lams.append(lam)
                kLs.append(k0 * L["thickness"])
            Sl = smatrix_layer(np.stack(Ws_), np.stack(Vs2), np.stack(lams), W0, V0, np.array(kLs))
            S = smatrix_star(S_ref, Sl) if S is None else smatrix_star(S, Sl)
            step += 1
            if Progress is not None:
                Progress(step / nsteps, f"layer {j + 1}/{nL}")
        if S is None:
            S = {k: np.broadcast_to(v, (len(ch),) + v.shape) for k, v in S_ref.items()}
        S = smatrix_star(S, S_trn)
        S11_all.append(S["S11"]); S21_all.append(S["S21"])
        if nL == 0 and Progress is not None:
            Progress(1.0, "done")
    S11 = np.concatenate(S11_all)
    S21 = np.concatenate(S21_all)

    i0 = int(np.flatnonzero((m == 0) & (n == 0))[0])
    c = np.zeros(2 * M, dtype=complex)
    c[i0] = E_inc[0]
    c[M + i0] = E_inc[1]
    if use_sym:
        rr = (S11 @ (UE.conj().T @ c)) @ UE.T
        tt = (S21 @ (UE.conj().T @ c)) @ UE.T
    else:
        rr = S11 @ c
        tt = S21 @ c
    kz_i = kz_nonzero(kz_branch(eps_inc - kx ** 2 - ky ** 2))
    kz_s = kz_nonzero(kz_branch(eps_sub - kx ** 2 - ky ** 2))
    rx, ry = rr[:, :M], rr[:, M:]
    tx, ty = tt[:, :M], tt[:, M:]
    rz = (kx * rx + ky * ry) / kz_i
    tz = -(kx * tx + ky * ty) / kz_s
    r = np.stack([rx, ry, rz], axis=-1)
    t = np.stack([tx, ty, tz], axis=-1)
    kzi0 = n_inc * ct
    R = kz_i.real / kzi0 * np.sum(np.abs(r) ** 2, axis=-1)
    T = kz_s.real / kzi0 * np.sum(np.abs(t) ** 2, axis=-1)
    return {"m": m, "n": n, "R": R, "T": T, "Rtot": R.sum(axis=1), "Ttot": T.sum(axis=1),
            "r": r, "t": t, "symmetry": use_sym, "Nsample": (Nsx, Nsy)}


def rcwa1d(dev, lambda0, theta, phi, pol, N=10, Nsample=0, Factorization="li",
           GapEps=float("nan")):
    """rcwa2D with Ny = 0 (rcwa1D.m:20-27)."""
    if not (math.isfinite(dev["period"][0]) and not math.isfinite(dev["period"][1])):
        raise OracleError("rcwa1D:period", "rcwa1D needs period [Lx Inf].")
    Ns = Nsample
    if Ns == 0:
        Ns = aligned_sample_count([dev], dev["period"], 1, max(1024, 8 * (4 * N + 1)))
    return rcwa2d(dev, lambda0, theta, phi, pol, Nx=N, Nsample=(Ns, 1),
                  Factorization=Factorization, GapEps=GapEps)


# ------------------------------------------------------------------ TMM
def tmm(lambda0, nidx, d, theta0, pol):
    """Abeles characteristic-matrix TMM (tmm.m:25-71)."""
    lambda0 = np.atleast_1d(np.asarray(lambda0, float))[:, None]
    nidx = np.atleast_2d(np.asarray(nidx, complex))
    if nidx.shape[0] == 1:
        nidx = np.repeat(nidx, lambda0.shape[0], axis=0)
    theta0 = np.atleast_1d(np.asarray(theta0, float))[None, :]
    Mn = nidx.shape[1]
    if np.any(np.abs(nidx[:, 0].imag) >= 1e-12):
        raise OracleError("tmm:lossyIncident", "Incident medium must be lossless.")
    k0 = 2 * np.pi / lambda0
    kx = k0 * nidx[:, :1].real * np.sin(theta0)

    def adm(nj):
        kz = kz_branch(nj ** 2 * k0 ** 2 - kx ** 2)
        q = kz / k0 if pol == "s" else kz / (k0 * nj ** 2)
        return q, kz
    q0, _ = adm(nidx[:, :1])
    qs, _ = adm(nidx[:, Mn - 1:Mn])
    m11 = np.ones(kx.shape, complex); m12 = np.zeros(kx.shape, complex)
    m21 = np.zeros(kx.shape, complex); m22 = np.ones(kx.shape, complex)
    d = np.atleast_1d(np.asarray(d, float))
    for j in range(1, Mn - 1):
        q, kz = adm(nidx[:, j:j + 1])
        delta = kz * d[j - 1]
        c, s = np.cos(delta), np.sin(delta)
        l11, l12, l21, l22 = c, -1j * s / q, -1j * q * s, c
        m11, m12, m21, m22 = (m11 * l11 + m12 * l21, m11 * l12 + m12 * l22,
                              m21 * l11 + m22 * l21, m21 * l12 + m22 * l22)
    Bq = m11 + m12 * qs
    Cq = m21 + m22 * qs
    den = q0 * Bq + Cq
    r = (q0 * Bq - Cq) / den
    t = 2 * q0 / den
    R = np.abs(r) ** 2
    T = qs.real / q0.real * np.abs(t) ** 2
    return {"r": r, "t": t, "R": R, "T": T, "A": 1 - R - T}


# ------------------------------------------------------------------ callers
def build_phase_library(makeCell: Callable, params, lambda0, N=10, Pol="s", CheckN=0,
                        Symmetry="auto"):
    """Sweep -> one batched rcwa2D; co-pol zero-order t (buildPhaseLibrary.m:2-17)."""
    devs = [makeCell(p) for p in params]
    def run(NN):
        o = rcwa2d(devs, lambda0, 0.0, 0.0, Pol, Nx=NN, Symmetry=Symmetry)
        k0 = int(np.flatnonzero((o["m"] == 0) & (o["n"] == 0))[0])
        comp = 1 if Pol == "s" else 0
        return o["t"][:, k0, comp], o["T"][:, k0]
    t, T = run(N)
    phase = np.unwrap(np.angle(t / t[0]))
    lib = {"param": np.asarray(params), "t": t, "T": T, "phase": phase,
           "span": float(phase.max() - phase.min()),
           "meta": {"lambda0": lambda0, "N": N, "pol": Pol, "period": devs[0]["period"]}}
    if CheckN > 0:
        t2, T2 = run(CheckN)
        dphi = np.abs(np.angle(t2 / t))
        lib["meta"].update({"CheckN": CheckN, "phaseDiff": dphi, "maxPhaseDiff": float(dphi.max()),
                            "maxTDiff": float(np.max(np.abs(T2 - T)))})
    return lib


class OracleWarning(UserWarning):
    """Warning carrying the reference's MATLAB warning identifier."""

    def __init__(self, ident: str, msg: str):
        super().__init__(f"{ident}: {msg}")
        self.identifier = ident


def converge_n(fun: Callable, Tol=1e-3, Ns=range(3, 26, 2), Progress=None):
    """Raise N until successive results differ by < Tol (convergeN.m:20-44).

    ``history`` mirrors the MATLAB table (convergeN.m:40): columns N, change,
    seconds over the N values tried; change is NaN for the first one.  A
    ladder that ends unconverged warns 'orf:convergeN:notConverged'
    (convergeN.m:41-44).
    """
    import time
    import warnings
    Ns = list(Ns)
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
        warnings.warn(OracleWarning("orf:convergeN:notConverged",
                                    f"Not converged to {Tol:g} by N = {res['N']} "
                                    f"(last change {hist['change'][-1]:.2g})."))
    return res


# ------------------------------------------------------------------ fields
def _star2(SA, SB):
    """Star product of two unbatched S-matrix dicts (smatrixStar.m:4-10)."""
    return {k: v[0] if v.ndim == 3 else v for k, v in smatrix_star(SA, SB).items()}


def rcwa_fields(dev, lambda0, theta, phi, pol, x, y, z, Nx=5, Ny=None, Nsample=0, Factorization="li",
                GapEps=float("nan")):
    """Fields in and around a grating (rcwaFields.m:1-21; algorithm restated).

    Every layer (and the half-spaces) is expanded in its own modes with
    s = [Ex; Ey] = W (e^{-lam k0 z} a+ + e^{-lam k0 (L - z)} a-) and
    u = [H~x; H~y] = V (-e^{-lam k0 z} a+ + e^{-lam k0 (L - z)} a-), z measured
    from the top of the layer (the convention of smatrixLayer.m:6-18).  The
    amplitudes come from the zero-thickness gaps between layers: with S_up the
    stack above a gap and S_down the stack below it, c+ = (I - S_up22
    S_down11)^-1 S_up21 c_inc and c- = S_down11 c+; a layer's a+ follows from
    the gap above it and a- from the gap below it, so no growing exponential
    is ever formed.  Ez = i [eps]^-1 (Kx H~y - Ky H~x) (pqMatrices.m:4),
    H~z = i (Kx Ey - Ky Ex).  Returns E, H (nx x ny x nz x 3), eps, flux
    (normalised to the incident flux) and the rcwa2D-style result.
    """
    x = np.atleast_1d(np.asarray(x, float))
    y = np.atleast_1d(np.asarray(y, float))
    z = np.atleast_1d(np.asarray(z, float))
    Lx, Ly = dev["period"]
    if Ny is None:
        Ny = Nx
    if not math.isfinite(Ly):
        Ny = 0
    m, n = harmonics2d(Nx, Ny)
    M = m.size
    k0 = 2 * np.pi / lambda0
    n_inc = nk(dev["incident"], lambda0)
    if abs(n_inc.imag) > 1e-12:
        raise OracleError("rcwa:lossyIncident", "Incident medium must be lossless.")
    n_inc = n_inc.real
    eps_inc = n_inc ** 2
    eps_sub = nk(dev["substrate"], lambda0) ** 2
    st, ct = math.sin(theta), math.cos(theta)
    khat = np.array([st * math.cos(phi), st * math.sin(phi), ct])
    E_inc = polarization_vector(khat, pol)
    kx = (n_inc * khat[0] + m * lambda0 / Lx).astype(float)
    ky = (n_inc * khat[1] + (n * lambda0 / Ly if math.isfinite(Ly) else 0.0 * n)).astype(float)
    if isinstance(Nsample, (int, float)) and Nsample == 0:
        Nsx, Nsy = default_sample_counts([dev], Nx, Ny)
    else:
        Nsx, Nsy = int(Nsample[0]), int(Nsample[1])
    xg, yg = cell_grid([Lx, Ly], Nsx, Nsy)
    eps_gap = GapEps if not math.isnan(GapEps) else 1.0 + float(np.max(kx ** 2 + ky ** 2))
    W0, V0, _, _ = modes_homogeneous(kx, ky, eps_gap)
    layers = []
    for L in dev["layers"]:
        if L["shapes"]:
            e = layer_eps(L, [Lx, Ly], xg, yg, lambda0)
            E = conv_matrix_2d(e, Nx, Ny)
            Exx, Eyy = conv_matrices_li_2d(e, Nx, Ny) if Factorization == "li" else (E, E)
            P, Q, Ei = pq_matrices(E, Exx, Eyy, kx, ky)
            W, V, lam = modes_layer(P, Q)
        else:
            eps = nk(L["background"], lambda0) ** 2
            W, V, lam, _ = modes_homogeneous(kx, ky, eps)
            Ei = np.eye(M) / eps
        S = smatrix_layer(W[None], V[None], lam[None], W0, V0, np.array([k0 * L["thickness"]]))
        layers.append({"W": W, "V": V, "lam": lam, "Ei": Ei, "d": L["thickness"],
                       "S": {k: v[0] for k, v in S.items()}})
    Wr, Vr, lamr, _ = modes_homogeneous(kx, ky, eps_inc)
    Ws, Vs, lams, _ = modes_homogeneous(kx, ky, eps_sub)
    S_ref = smatrix_region(Wr, Vr, W0, V0, "ref")
    S_trn = smatrix_region(Ws, Vs, W0, V0, "trn")
    nL = len(layers)
    # prefix stacks S_ref * S_1 * ... * S_{j}, suffix stacks S_{j+1} * ... * S_trn
    up = [S_ref]
    for j in range(nL):
        up.append(_star2(up[-1], layers[j]["S"]))
    down = [None] * (nL + 1)
    down[nL] = S_trn
    for j in range(nL - 1, -1, -1):
        down[j] = _star2(layers[j]["S"], down[j + 1])
    glob = _star2(up[nL], S_trn)
    i0 = int(np.flatnonzero((m == 0) & (n == 0))[0])
    c_inc = np.zeros(2 * M, complex)
    c_inc[i0], c_inc[M + i0] = E_inc[0], E_inc[1]
    c_ref = glob["S11"] @ c_inc
    c_trn = glob["S21"] @ c_inc
    I2 = np.eye(2 * M)
    gaps = []   # gap above layer j (j = 0..nL): (c+, c-)
    for j in range(nL + 1):
        cp = np.linalg.solve(I2 - up[j]["S22"] @ down[j]["S11"], up[j]["S21"] @ c_inc)
        gaps.append((cp, down[j]["S11"] @ cp))
    ztop = np.concatenate([[0.0], np.cumsum([Lj["d"] for Lj in layers])])
    nz = z.size
    s_all = np.zeros((nz, 2 * M), complex)
    u_all = np.zeros((nz, 2 * M), complex)
    ez_all = np.zeros((nz, M), complex)
    layer_of = np.full(nz, -1)
    for iz, zz in enumerate(z):
        if zz < 0:
            fp = np.exp(-lamr * k0 * zz)
            fm = np.exp(lamr * k0 * zz)
            s = c_inc * fp + c_ref * fm
            u = Vr @ (-c_inc * fp + c_ref * fm)
            Ei = np.eye(M) / eps_inc
        elif zz >= ztop[-1]:
            fp = np.exp(-lams * k0 * (zz - ztop[-1]))
            s = c_trn * fp
            u = Vs @ (-c_trn * fp)
            Ei = np.eye(M) / eps_sub
            layer_of[iz] = nL
        else:
            j = int(np.searchsorted(ztop, zz, side="right") - 1)
            Lj = layers[j]
            layer_of[iz] = j
            cp, cm = gaps[j]
            dp, dm = gaps[j + 1]
            WiW0 = np.linalg.solve(Lj["W"], W0)
            ViV0 = np.linalg.solve(Lj["V"], V0)
            ap = 0.5 * (WiW0 @ (cp + cm) + ViV0 @ (cp - cm))
            am = 0.5 * (WiW0 @ (dp + dm) + ViV0 @ (dm - dp))
            zl = zz - ztop[j]
            fp = np.exp(-Lj["lam"] * k0 * zl)
            fm = np.exp(-Lj["lam"] * k0 * (Lj["d"] - zl))
            s = Lj["W"] @ (fp * ap + fm * am)
            u = Lj["V"] @ (-fp * ap + fm * am)
            Ei = Lj["Ei"]
        s_all[iz], u_all[iz] = s, u
        ez_all[iz] = 1j * Ei @ (kx * u[M:] - ky * u[:M])
    hz_all = 1j * (kx * s_all[:, M:] - ky * s_all[:, :M])
    # synthesis on the (x, y) grid: phase exp(i k0 (kx x + ky y))
    ph = np.exp(1j * k0 * (kx[None, None, :] * x[:, None, None] + ky[None, None, :] * y[None, :, None]))
    def synth(c):   # c: nz x M -> nx x ny x nz
        return np.einsum("abp,zp->abz", ph, c)
    E = np.stack([synth(s_all[:, :M]), synth(s_all[:, M:]), synth(ez_all)], axis=-1)
    H = np.stack([synth(u_all[:, :M]), synth(u_all[:, M:]), synth(hz_all)], axis=-1)
    flux = -np.imag(np.sum(s_all[:, :M] * np.conj(u_all[:, M:]) - s_all[:, M:] * np.conj(u_all[:, :M]),
                           axis=1)) / (n_inc * ct)
    # permittivity at the field points
    eps = np.zeros((x.size, y.size, nz), complex)
    for iz in range(nz):
        j = layer_of[iz]
        if j < 0:
            eps[:, :, iz] = eps_inc
        elif j == nL:
            eps[:, :, iz] = eps_sub
        else:
            eps[:, :, iz] = layer_eps(dev["layers"][j], [Lx, Ly], x, y, lambda0)
    out = rcwa2d(dev, lambda0, theta, phi, pol, Nx=Nx, Ny=Ny, Nsample=(Nsx, Nsy),
                 Factorization=Factorization, GapEps=GapEps, Symmetry="none")
    return {"E": E, "H": H, "eps": eps, "x": x, "y": y, "z": z, "zInterfaces": ztop, "flux": flux,
            "out": out, "coeffs": {"s": s_all, "u": u_all, "ez": ez_all, "hz": hz_all}}
