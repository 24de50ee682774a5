"""Pins of the FP64 CPU oracle (oracle/rcwa_oracle.py).

Each test ports one of the reference's own RCWA tests (file:line given) or a
closed form the mathematics fixes, so that a plausible mistake in the
restatement (a sign of kx, a transposed Toeplitz, the wrong Li direction, a
dropped z-component) fails at least one of them.
"""
import json
import math
import os

import numpy as np
import pytest

from oracle import rcwa_oracle as O
from opticrcwafdtd_amd import structure as S

mc = S.materialConst
GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reticolo_v10.json")))
TOL = GOLD["tolerance"]


def _idx(o, m, n=0):
    return int(np.flatnonzero((o["m"] == m) & (o["n"] == n))[0])


# ------------------------------------------------------------------ RETICOLO (tests/TestReticolo.m)
def test_reticolo_silicon_lines_1d():
    g = S.makeStructure([0.7, math.inf], mc(1), mc(1.45),
                        [S.makeLayer(0.5, mc(1), [S.shapeRect(0.35, math.inf, mc(3.48))])])
    for pol in "sp":
        o = O.rcwa1d(g, 1.0, math.pi / 6, 0, pol, N=20)
        k = [_idx(o, -1), _idx(o, 0)]
        got = np.array([o["R"][0, k], o["T"][0, k]])
        np.testing.assert_allclose(got, GOLD["silicon_lines_1d"][pol], atol=TOL, rtol=0)


def test_reticolo_metal_lines_1d():
    nAg = np.sqrt(-40 + 3j)
    g = S.makeStructure([0.5, math.inf], mc(1), mc(1.45),
                        [S.makeLayer(0.1, mc(1), [S.shapeRect(0.25, math.inf, mc(nAg))])])
    o = O.rcwa1d(g, 0.8, 10 * math.pi / 180, 0, "p", N=10)
    G = GOLD["metal_lines_1d"]
    assert abs(o["R"][0, _idx(o, 0)] - G["R0"]) < TOL
    assert abs(o["T"][0, _idx(o, -1)] - G["T_m1"]) < TOL
    assert abs(o["T"][0, _idx(o, 0)] - G["T_0"]) < TOL


def asymmetric_cell():
    m2 = mc(2)
    return S.makeStructure([0.6, 0.5], mc(1), mc(1.45),
                           [S.makeLayer(0.3, mc(1), [S.shapeRect(0.2, 0.3, m2, [-0.15, 0.05]),
                                                     S.shapeRect(0.1, 0.15, m2, [0.15, -0.1])])])


def test_reticolo_asymmetric_cell_2d():
    # Nsample: any edge-aligned grid gives the exact staircase for rectangles.
    o = O.rcwa2d(asymmetric_cell(), 0.55, 25 * math.pi / 180, 60 * math.pi / 180, "p", Nx=7,
                 Symmetry="none", Nsample=(120, 120))
    G = GOLD["asymmetric_cell_2d"]
    for m, n, v in G["R"]:
        assert abs(o["R"][0, _idx(o, m, n)] - v) < TOL
    for m, n, v in G["T"]:
        assert abs(o["T"][0, _idx(o, m, n)] - v) < TOL
    tot = sum(r[2] for r in G["R"]) + sum(r[2] for r in G["T"])
    assert abs(o["Rtot"][0] + o["Ttot"][0] - tot) < TOL


def square_pillars(ws):
    return [S.makeStructure([0.4, 0.4], mc(1.45), mc(1),
                            [S.makeLayer(0.5, mc(1), [S.shapeRect(a, a, mc(3.6))])]) for a in ws]


@pytest.mark.parametrize("sym", ["auto", "none"])
def test_reticolo_phase_library(sym):
    G = GOLD["phase_library"]
    o = O.rcwa2d(square_pillars(G["w"]), 0.94, 0, 0, "s", Nx=6, Symmetry=sym, Nsample=(80, 80))
    assert o["symmetry"] == (sym == "auto")
    k0 = _idx(o, 0, 0)
    t = o["t"][:, k0, 0] + o["t"][:, k0, 1]
    np.testing.assert_allclose(o["T"][:, k0], G["T0"], atol=TOL, rtol=0)
    ph = np.array(G["phase"])
    d = np.angle(t / t[0]) - np.angle(np.exp(1j * (ph - ph[0])))
    assert np.max(np.abs(d)) < GOLD["phase_tolerance"]


def _square_cell(a):
    return S.makeStructure([0.4, 0.4], mc(1.45), mc(1), [S.makeLayer(0.5, mc(1), [S.shapeRect(a, a, mc(3.6))])])


def test_build_phase_library_reticolo():
    """tests/TestReticolo.m:63-78 through the phase-library caller (buildPhaseLibrary.m:2-17):
    co-polarised zero-order T and the phase relative to the first cell."""
    G = GOLD["phase_library"]
    lib = O.build_phase_library(_square_cell, G["w"], 0.94, N=6, Pol="s")
    np.testing.assert_allclose(lib["T"], G["T0"], atol=TOL, rtol=0)
    ph = np.array(G["phase"])
    assert lib["phase"][0] == 0.0
    d = np.angle(np.exp(1j * lib["phase"])) - np.angle(np.exp(1j * (ph - ph[0])))
    assert np.max(np.abs(d)) < GOLD["phase_tolerance"]
    # unwrapped: no jump between neighbouring cells exceeds pi; span is the coverage
    assert np.all(np.abs(np.diff(lib["phase"])) <= math.pi)
    assert abs(lib["span"] - (lib["phase"].max() - lib["phase"].min())) == 0.0
    assert lib["meta"]["N"] == 6 and lib["meta"]["pol"] == "s" and list(lib["meta"]["period"]) == [0.4, 0.4]


def test_build_phase_library_checkn():
    """CheckN (buildPhaseLibrary.m:7-9): the rerun at CheckN = N changes nothing, and at
    another N the stored differences are those of the pinned rcwa2d at that N."""
    w = [0.1, 0.2, 0.3]
    same = O.build_phase_library(_square_cell, w, 0.94, N=4, CheckN=4)
    assert same["meta"]["maxPhaseDiff"] == 0.0 and same["meta"]["maxTDiff"] == 0.0
    lib = O.build_phase_library(_square_cell, w, 0.94, N=4, CheckN=5)
    o5 = O.rcwa2d([_square_cell(a) for a in w], 0.94, 0, 0, "s", Nx=5)
    k0 = _idx(o5, 0, 0)
    np.testing.assert_allclose(lib["meta"]["maxTDiff"], np.max(np.abs(o5["T"][:, k0] - lib["T"])), atol=1e-15)
    assert np.all(lib["meta"]["phaseDiff"] >= 0) and lib["meta"]["maxPhaseDiff"] < 0.2
    # p-pol: the co-polarised component is t_x; the square pillar is polarisation independent
    lp = O.build_phase_library(_square_cell, w, 0.94, N=4, Pol="p")
    np.testing.assert_allclose(lp["T"], lib["T"], atol=1e-12)
    np.testing.assert_allclose(lp["phase"], lib["phase"], atol=1e-11)


def test_converge_n_ladder_logic():
    """convergeN.m:20-44 on closed-form sequences: the first N whose change is below Tol,
    NaN change for the first entry, Progress fractions (k-1)/numel(Ns)."""
    log = []
    r = O.converge_n(lambda N: 2.0 ** -N, Tol=0.1, Ns=[1, 2, 3, 4, 5, 6],
                     Progress=lambda f, msg: log.append((f, msg)))
    assert r["converged"] and r["N"] == 4 and r["value"][0] == 1 / 16
    assert r["history"]["N"] == [1, 2, 3, 4]
    ch = r["history"]["change"]
    assert math.isnan(ch[0]) and ch[1:] == [0.25, 0.125, 0.0625]
    assert [f for f, _ in log] == [0, 1 / 6, 2 / 6, 3 / 6] and log[2][1] == "N = 3"
    # vector results: max abs over the elements, complex included
    r = O.converge_n(lambda N: np.array([1.0, 1j / N]), Tol=0.3, Ns=[1, 2, 3])
    assert r["N"] == 3 and abs(r["history"]["change"][1] - 0.5) < 1e-15


def test_converge_n_warns_when_not_converged():
    """tests/TestGUISupport.m:66-67: fun = 1/N, Tol = 1e-9, Ns = 1:3 warns
    'orf:convergeN:notConverged' and reports the last N."""
    with pytest.warns(O.OracleWarning) as rec:
        r = O.converge_n(lambda N: 1 / N, Tol=1e-9, Ns=[1, 2, 3])
    assert rec[0].message.identifier == "orf:convergeN:notConverged"
    assert not r["converged"] and r["N"] == 3 and abs(r["value"][0] - 1 / 3) < 1e-16
    assert r["history"]["N"] == [1, 2, 3]


@pytest.mark.slow
def test_converge_n_two_layer_cell():
    """tests/TestGUISupport.m:60-65: Ttot of the two-layer cell converges to 1e-3 on
    Ns = 1:2:15, the last change is below Tol, and the value equals a direct solve."""
    dev = two_layers()
    r = O.converge_n(lambda N: O.rcwa2d(dev, 1.0, 0, 0, "s", Nx=N)["Ttot"], Tol=1e-3, Ns=range(1, 16, 2))
    assert r["converged"] and r["history"]["change"][-1] < 1e-3
    np.testing.assert_array_equal(r["value"], O.rcwa2d(dev, 1.0, 0, 0, "s", Nx=r["N"])["Ttot"])


# ------------------------------------------------------------------ TMM closed forms (tests/TestTMM.m)
def test_tmm_fresnel_and_energy():
    n1, n2 = 1.0, 1.5
    th1 = np.deg2rad(np.arange(0, 86, 5))
    th2 = np.arcsin(n1 * np.sin(th1) / n2)
    rs = (n1 * np.cos(th1) - n2 * np.cos(th2)) / (n1 * np.cos(th1) + n2 * np.cos(th2))
    rp = (n2 * np.cos(th1) - n1 * np.cos(th2)) / (n2 * np.cos(th1) + n1 * np.cos(th2))
    s = O.tmm(0.633, [n1, n2], [], th1, "s")
    p = O.tmm(0.633, [n1, n2], [], th1, "p")
    np.testing.assert_allclose(s["r"][0], rs, atol=1e-12)
    np.testing.assert_allclose(p["r"][0], rp, atol=1e-12)
    np.testing.assert_allclose(s["R"] + s["T"], 1, atol=1e-12)
    assert O.tmm(0.633, [1, 1.5], [], math.atan(1.5), "p")["R"][0, 0] < 1e-28


def test_tmm_airy_slab_with_loss_and_bragg():
    lam = np.linspace(0.5, 1.5, 41)
    n = 3.5 + 0.05j
    d = 0.4
    delta = 2 * np.pi / lam * n * d
    r12 = (1 - n) / (1 + n)
    rref = r12 * (1 - np.exp(2j * delta)) / (1 - r12 ** 2 * np.exp(2j * delta))
    o = O.tmm(lam, [1, n, 1], [d], 0, "s")
    np.testing.assert_allclose(o["r"][:, 0], rref, atol=1e-12)
    assert np.all(o["A"] > 0)
    lam0, nH, nL, N = 1.55, 3.48, 1.45, 6
    nn = [1] + [nH, nL] * N + [1.45]
    dd = [lam0 / (4 * nH), lam0 / (4 * nL)] * N
    o = O.tmm(lam0, nn, dd, 0, "s")
    x, y = nL ** (2 * N), 1.45 * nH ** (2 * N)
    assert abs(o["R"][0, 0] - ((x - y) / (x + y)) ** 2) < 1e-12


def test_tmm_total_internal_reflection():
    th = np.deg2rad([45, 60, 80])
    for pol in "sp":
        o = O.tmm(0.633, [1.5, 1.0], [], th, pol)
        np.testing.assert_allclose(o["R"], 1, atol=1e-12)
        np.testing.assert_allclose(o["T"], 0, atol=1e-12)


# ------------------------------------------------------------------ RCWA vs TMM
def test_homogeneous_conical_equals_tmm():
    """tests/TestRCWA2D.m:25-33."""
    th, ph = math.radians(30), math.radians(40)
    dev = S.makeStructure([0.5, 0.4], mc(1), mc(3.48),
                          [S.makeLayer(0.2, mc(1.45)), S.makeLayer(0.11, mc(2 + 0.1j))])
    for pol in "sp":
        o = O.rcwa2d(dev, 1.3, th, ph, pol, Nx=3)
        t = O.tmm(1.3, [1, 1.45, 2 + 0.1j, 3.48], [0.2, 0.11], th, pol)
        assert abs(o["Rtot"][0] - t["R"][0, 0]) < 1e-12
        assert abs(o["Ttot"][0] - t["T"][0, 0]) < 1e-12


def test_homogeneous_amplitudes_equal_tmm():
    """tests/TestRCWA1D.m:31-47: r, t amplitudes (phase reference) and p-pol R, T."""
    dev = S.makeStructure([0.6, math.inf], mc(1), mc(3.48),
                          [S.makeLayer(0.2, mc(1.45)), S.makeLayer(0.11, mc(2 + 0.1j))])
    nn, d = [1, 1.45, 2 + 0.1j, 3.48], [0.2, 0.11]
    for th in np.deg2rad([0, 25, 60]):
        o = O.rcwa1d(dev, 1.3, th, 0, "s", N=5)
        t = O.tmm(1.3, nn, d, th, "s")
        k = _idx(o, 0)
        assert abs(o["r"][0, k, 1] - t["r"][0, 0]) < 1e-12
        assert abs(o["t"][0, k, 1] - t["t"][0, 0]) < 1e-12
        o = O.rcwa1d(dev, 1.3, th, 0, "p", N=5)
        t = O.tmm(1.3, nn, d, th, "p")
        assert abs(o["Rtot"][0] - t["R"][0, 0]) < 1e-12
        assert abs(o["Ttot"][0] - t["T"][0, 0]) < 1e-12


def test_eig_path_equals_tmm():
    """tests/TestRCWA1D.m:49-58: a patterned-but-uniform layer through eig."""
    sio2 = mc(1.45)
    dev = S.makeStructure([0.6, math.inf], mc(1), mc(3.48),
                          [S.makeLayer(0.2, sio2, [S.shapeRect(0.3, math.inf, sio2)])])
    t = O.tmm(1.3, [1, 1.45, 3.48], [0.2], math.radians(25), "p")
    o = O.rcwa1d(dev, 1.3, math.radians(25), 0, "p", N=5)
    assert abs(o["Rtot"][0] - t["R"][0, 0]) < 1e-12
    assert abs(o["Ttot"][0] - t["T"][0, 0]) < 1e-12


def test_conv_matrix_exact_for_aligned_rect():
    """tests/TestRCWA1D.m:19-29."""
    Nx, f, e1, e2, N = 64, 0.25, 4.0, 1.0, 5
    eps = e2 * np.ones(Nx)
    eps[Nx // 2 - int(Nx * f / 2): Nx // 2 + int(Nx * f / 2)] = e1
    C = O.conv_matrix_1d(eps, N)
    h = np.arange(0, 2 * N + 1)
    ref = (e1 - e2) * f * np.sinc(h * f) + e2 * (h == 0)
    np.testing.assert_allclose(C[:, 0], ref, atol=1e-14)


def test_li_2d_reduces_to_1d():
    """tests/TestRCWA2D.m:35-47: a y-invariant line in a finite-Ly lattice."""
    th, ph = math.radians(30), math.radians(40)
    air, si = mc(1), mc(3.48)
    g1 = S.makeStructure([0.7, math.inf], air, mc(1.45), [S.makeLayer(0.5, air, [S.shapeRect(0.35, math.inf, si)])])
    g2 = S.makeStructure([0.7, 0.5], air, mc(1.45), [S.makeLayer(0.5, air, [S.shapeRect(0.35, math.inf, si)])])
    a = O.rcwa1d(g1, 1.0, th, ph, [1, 1j], N=10, Nsample=256)
    b = O.rcwa2d(g2, 1.0, th, ph, [1, 1j], Nx=10, Ny=3, Nsample=(256, 64))
    assert abs(b["Rtot"][0] - a["Rtot"][0]) < 1e-12
    assert abs(b["Ttot"][0] - a["Ttot"][0]) < 1e-12


# ------------------------------------------------------------------ invariants (tests/TestRCWA2D.m)
def pillar_sweep(Ds, nsi=3.48):
    return [S.makeStructure([0.4, 0.4], mc(1.45), mc(1),
                            [S.makeLayer(0.6, mc(1), [S.shapeCircle(D / 2, mc(nsi))])]) for D in Ds]


def test_rotation_swaps_polarization():
    sio2, air, si = mc(1.45), mc(1), mc(3.48)
    r1 = S.makeStructure([0.5, 0.5], sio2, air, [S.makeLayer(0.6, air, [S.shapeRect(0.3, 0.15, si)])])
    r2 = S.makeStructure([0.5, 0.5], sio2, air, [S.makeLayer(0.6, air, [S.shapeRect(0.15, 0.3, si)])])
    a = O.rcwa2d(r1, 1.0, 0, 0, "s", Nx=6, Nsample=(100, 100))
    b = O.rcwa2d(r2, 1.0, 0, 0, "p", Nx=6, Nsample=(100, 100))
    assert abs(a["Ttot"][0] - b["Ttot"][0]) < 1e-12
    assert abs(a["Rtot"][0] - b["Rtot"][0]) < 1e-12


def test_circle_is_polarization_independent():
    devs = pillar_sweep([0.24])
    a = O.rcwa2d(devs, 0.94, 0, 0, "s", Nx=6)
    b = O.rcwa2d(devs, 0.94, 0, 0, "p", Nx=6)
    assert abs(a["Ttot"][0] - b["Ttot"][0]) < 1e-12
    k = _idx(a, 0, 0)
    assert abs(abs(a["t"][0, k, 1]) - abs(b["t"][0, k, 0])) < 1e-12


def test_energy_conservation_conical():
    c = S.makeStructure([0.6, 0.6], mc(1), mc(1.45), [S.makeLayer(0.5, mc(1), [S.shapeCircle(0.18, mc(3.48))])])
    o = O.rcwa2d(c, 1.0, math.radians(30), math.radians(40), [1, 1j], Nx=6)
    assert abs(o["Rtot"][0] + o["Ttot"][0] - 1) < 1e-12


def test_batch_equals_loop_and_chunking():
    devs = pillar_sweep([0.1, 0.15, 0.2, 0.25, 0.3])
    ob = O.rcwa2d(devs, 0.94, 0, 0, "s", Nx=3)
    oc = O.rcwa2d(devs, 0.94, 0, 0, "s", Nx=3, BatchSize=2)
    np.testing.assert_allclose(oc["t"], ob["t"], atol=1e-12)
    for k in (0, 2, 4):
        ol = O.rcwa2d(devs[k], 0.94, 0, 0, "s", Nx=3)
        np.testing.assert_allclose(ob["t"][k], ol["t"][0], atol=1e-12)


def test_symmetry_reduction_matches_full():
    """tests/TestRCWA2D.m:87-102 (with a homogeneous layer in the stack)."""
    devs = [S.makeStructure([0.4, 0.4], mc(1.45), mc(1),
                            [S.makeLayer(0.1, mc(2.0)),
                             S.makeLayer(0.6, mc(1), [S.shapeCircle(D / 2, mc(3.48))])]) for D in (0.15, 0.3)]
    for pol in "sp":
        a = O.rcwa2d(devs, 0.94, 0, 0, pol, Nx=5)
        b = O.rcwa2d(devs, 0.94, 0, 0, pol, Nx=5, Symmetry="none")
        assert a["symmetry"] and not b["symmetry"]
        np.testing.assert_allclose(a["t"], b["t"], atol=1e-11)
        np.testing.assert_allclose(a["r"], b["r"], atol=1e-11)
        np.testing.assert_allclose(a["T"], b["T"], atol=1e-11)


def test_asymmetric_falls_back_to_full():
    dev = S.makeStructure([0.5, 0.5], mc(1), mc(1.45),
                          [S.makeLayer(0.4, mc(1), [S.shapeCircle(0.1, mc(2.5), [0.07, 0])])])
    a = O.rcwa2d(dev, 1.0, 0, 0, "p", Nx=4)
    b = O.rcwa2d(dev, 1.0, 0, 0, "p", Nx=4, Symmetry="none")
    assert not a["symmetry"]
    np.testing.assert_allclose(a["t"], b["t"], atol=1e-12)
    assert abs(a["Rtot"][0] + a["Ttot"][0] - 1) < 1e-12


# ------------------------------------------------------------------ 1D invariants (tests/TestRCWA1D.m)
def si_grating():
    return S.makeStructure([0.7, math.inf], mc(1), mc(1.45),
                           [S.makeLayer(0.5, mc(1), [S.shapeRect(0.35, math.inf, mc(3.48))])])


def test_energy_gap_mirror_subwavelength():
    g = si_grating()
    for th, ph, pol in [(0, 0, "s"), (math.radians(20), 0, "s"), (math.radians(20), 0, "p"),
                        (math.radians(20), math.radians(30), [1, 1j])]:
        o = O.rcwa1d(g, 1.0, th, ph, pol, N=15)
        assert abs(o["Rtot"][0] + o["Ttot"][0] - 1) < 1e-12
    a = O.rcwa1d(g, 1.0, math.radians(20), 0, "p", N=10)
    b = O.rcwa1d(g, 1.0, math.radians(20), 0, "p", N=10, GapEps=2.3)
    np.testing.assert_allclose(a["r"], b["r"], atol=1e-12)
    np.testing.assert_allclose(a["t"], b["t"], atol=1e-12)
    a = O.rcwa1d(g, 1.0, math.radians(20), 0, "s", N=10)
    b = O.rcwa1d(g, 1.0, math.radians(-20), 0, "s", N=10)
    np.testing.assert_allclose(a["R"][0], b["R"][0][::-1], atol=1e-12)
    np.testing.assert_allclose(a["T"][0], b["T"][0][::-1], atol=1e-12)
    o = O.rcwa1d(g, 1.55, 0, 0, "s", N=10)
    nz = o["m"] != 0
    assert np.max(np.abs(o["R"][0, nz])) < 1e-14 and np.max(np.abs(o["T"][0, nz])) < 1e-14


def test_effective_medium_limit():
    P, e1, f = 0.01, 4.0, 0.5
    g = S.makeStructure([P, math.inf], mc(1), mc(1.45),
                        [S.makeLayer(0.3, mc(1), [S.shapeRect(f * P, math.inf, mc(math.sqrt(e1)))])])
    epar = f * e1 + (1 - f)
    eperp = 1 / (f / e1 + (1 - f))
    s = O.rcwa1d(g, 1.0, 0, 0, "s", N=10)
    p = O.rcwa1d(g, 1.0, 0, 0, "p", N=10)
    assert abs(s["Rtot"][0] - O.tmm(1.0, [1, math.sqrt(epar), 1.45], [0.3], 0, "s")["R"][0, 0]) < 1e-4
    assert abs(p["Rtot"][0] - O.tmm(1.0, [1, math.sqrt(eperp), 1.45], [0.3], 0, "p")["R"][0, 0]) < 1e-3


def test_li_converges_faster_than_laurent():
    g = si_grating()
    ref = O.rcwa1d(g, 1.55, 0, 0, "p", N=100, Nsample=4096)["T"][0, 100]
    li = O.rcwa1d(g, 1.55, 0, 0, "p", N=20, Nsample=4096)["T"][0, 20]
    la = O.rcwa1d(g, 1.55, 0, 0, "p", N=20, Nsample=4096, Factorization="laurent")["T"][0, 20]
    assert abs(li - ref) < 1e-4
    assert abs(li - ref) < abs(la - ref) / 10


def test_rayleigh_anomaly_stays_finite():
    g = si_grating()
    lam = 0.7 * 1.45
    for pol in "sp":
        o = O.rcwa1d(g, lam, 0, 0, pol, N=15)
        assert np.all(np.isfinite(o["R"])) and np.all(np.isfinite(o["T"]))
        assert abs(o["Rtot"][0] + o["Ttot"][0] - 1) < 1e-6
        oL = O.rcwa1d(g, lam - 1e-5, 0, 0, pol, N=15)
        oR = O.rcwa1d(g, lam + 1e-5, 0, 0, pol, N=15)
        assert abs(o["Ttot"][0] - (oL["Ttot"][0] + oR["Ttot"][0]) / 2) < 1e-2


def test_literature_lalanne_hugonin():
    """tests/TestLiteratureRCWA.m:20-28 (TM, N = 80)."""
    g = S.makeStructure([2, math.inf], mc(1), mc(1.5),
                        [S.makeLayer(1, mc(1), [S.shapeRect(0.468, math.inf, mc(2.3))])])
    o = O.rcwa1d(g, 1, math.pi / 6, 0, "p", N=80)
    assert abs(o["T"][0, _idx(o, 1)] - GOLD["literature"]["lalanne_hugonin_T1"]) < GOLD["literature"]["lh_tol"]


def test_literature_granet_te():
    """tests/TestLiteratureRCWA.m:12-18 (metallic lamellar grating, TE, N = 80)."""
    met = mc(0.22 + 6.71j)
    g = S.makeStructure([1, math.inf], mc(1), met, [S.makeLayer(1, mc(1), [S.shapeRect(0.5, math.inf, met)])])
    o = O.rcwa1d(g, 1, math.pi / 6, 0, "s", N=80)
    assert abs(o["R"][0, _idx(o, -1)] - GOLD["literature"]["granet_te_Rm1"]) < GOLD["literature"]["granet_tol"]


# ------------------------------------------------------------------ geometry (tests/TestPolygonsAndImport.m)
def test_polygon_equals_rect_and_vertical_sidewall():
    si = mc(3.5)
    sq = np.array([[-0.1, -0.1], [0.1, -0.1], [0.1, 0.1], [-0.1, 0.1]])
    mk = lambda s: S.makeStructure([0.4, 0.4], mc(1), mc(1.45), [S.makeLayer(0.3, mc(1), [s])])
    run = lambda s: O.rcwa2d(mk(s), 1.0, 0.2, 0.5, "p", Nx=4, Nsample=(400, 400), Symmetry="none")
    a = run(S.shapeRect(0.2, 0.2, si))
    b = run(S.shapePolygon(sq, si))
    c = run(S.shapePolygon(sq @ np.array([[0, 1], [-1, 0]]), si))
    np.testing.assert_allclose(b["T"], a["T"], atol=1e-12)
    np.testing.assert_allclose(c["T"], a["T"], atol=1e-12)
    air = mc(1)
    shp = S.shapeCircle(0.12, si)
    one = S.makeStructure([0.4, 0.4], air, mc(1.45), [S.makeLayer(0.4, air, [shp])])
    sl = S.makeStructure([0.4, 0.4], air, mc(1.45), S.slantedLayers(0.4, air, shp, 90, NSlices=4))
    a = O.rcwa2d(one, 1.0, 0, 0, "s", Nx=4)
    b = O.rcwa2d(sl, 1.0, 0, 0, "s", Nx=4)
    np.testing.assert_allclose(b["T"], a["T"], atol=1e-10)


def test_slanted_geometry():
    air, si = mc(1), mc(3.5)
    L = S.slantedLayers(0.4, air, S.shapeCircle(0.15, si), 80, NSlices=4)
    r = np.array([l["shapes"][0]["radius"] for l in L])
    cot80 = 1 / math.tan(math.radians(80))
    np.testing.assert_allclose(r, 0.15 - (0.4 - (np.arange(1, 5) - 0.5) * 0.1) * cot80, atol=1e-14)
    Lp = S.slantedLayers(0.4, air, S.shapePolygon([[-0.1, -0.1], [0.1, -0.1], [0.1, 0.1], [-0.1, 0.1]], si),
                         80, NSlices=2, Base="top")
    v = Lp[0]["shapes"][0]["vertices"]
    area = 0.5 * abs(np.sum(v[:, 0] * np.roll(v[:, 1], -1) - np.roll(v[:, 0], -1) * v[:, 1]))
    w = 0.2 - 2 * 0.1 * cot80
    assert abs(area - w * w) < 1e-8


# ------------------------------------------------------------------ progress / callers (tests/TestGUISupport.m)
def two_layers():
    si, air = mc(3.5), mc(1)
    return S.makeStructure([0.5, 0.5], air, mc(1.45),
                           [S.makeLayer(0.2, air, [S.shapeCircle(0.15, si)]),
                            S.makeLayer(0.1, air, [S.shapeRect(0.3, 0.2, si)])])


def test_progress_and_cancel():
    log = []
    O.rcwa2d(two_layers(), 1.0, 0, 0, "s", Nx=3, Progress=lambda f, msg: log.append(f))
    np.testing.assert_allclose(log, [0.5, 1.0], atol=1e-12)
    log = []
    O.rcwa2d([two_layers()] * 5, 1.0, 0, 0, "s", Nx=3, BatchSize=2, Progress=lambda f, msg: log.append(f))
    assert np.all(np.diff(log) > 0) and abs(log[-1] - 1) < 1e-12

    class Cancelled(Exception):
        pass

    def stop(f, msg):
        raise Cancelled("orf:cancelled")
    with pytest.raises(Cancelled):
        O.rcwa2d(two_layers(), 1.0, 0, 0, "s", Nx=3, Progress=stop)


def test_invalid_pol_raises():
    with pytest.raises(O.OracleError) as e:
        O.rcwa2d(two_layers(), 1.0, 0, 0, "x", Nx=1)
    assert e.value.identifier == "rcwa:pol"


# ------------------------------------------------------------------ fields (tests/TestRCWAFields.m)
def _field_pillars():
    return S.makeStructure([0.4, 0.4], mc(1.45), mc(1),
                           [S.makeLayer(0.5, mc(1), [S.shapeCircle(0.1, mc(3.6))]), S.makeLayer(0.2, mc(2))])


def test_fields_slab_matches_analytic_field():
    """tests/TestRCWAFields.m:11-31: E_y(z) of a slab built from the substrate upward."""
    nidx = [1, 2, 1.45]
    d, k = 0.3, 2 * np.pi
    slab = S.makeStructure([0.37, 0.37], mc(nidx[0]), mc(nidx[2]), [S.makeLayer(d, mc(nidx[1]))])
    z = np.linspace(-0.4, 0.7, 23)
    f = O.rcwa_fields(slab, 1, 0, 0, "s", 0, 0, z, Nx=1)
    Ey = f["E"][0, 0, :, 1]
    t = Ey[-1] / np.exp(1j * k * nidx[2] * (z[-1] - d))
    q = k * nidx[1]
    in_slab = lambda zz: t * np.cos(q * (zz - d)) + 1j * k * nidx[2] * t / q * np.sin(q * (zz - d))
    E0 = in_slab(0)
    dE0 = -q * t * np.sin(-q * d) + 1j * k * nidx[2] * t * np.cos(-q * d)
    a = (E0 + dE0 / (1j * k * nidx[0])) / 2
    b = (E0 - dE0 / (1j * k * nidx[0])) / 2
    ref = (z < 0) * (a * np.exp(1j * k * nidx[0] * z) + b * np.exp(-1j * k * nidx[0] * z)) + \
        ((z >= 0) & (z < d)) * in_slab(z) + (z >= d) * t * np.exp(1j * k * nidx[2] * (z - d))
    np.testing.assert_allclose(Ey, ref, atol=1e-12, rtol=0)
    assert abs(a - 1) < 1e-12
    assert abs(nidx[2] * abs(t) ** 2 - O.tmm(1, nidx, [d], 0, "s")["T"][0, 0]) < 1e-12


def test_fields_flux_is_conserved_and_equals_T():
    """tests/TestRCWAFields.m:33-44 (Nx = 3 here to keep the CPU suite short)."""
    dev = _field_pillars()
    z = np.array([-0.3, 0, 0.1, 0.25, 0.499, 0.5, 0.6, 0.69, 0.7, 1.5])
    for pol in "sp":
        o = O.rcwa2d(dev, 0.94, 0.3, 0.4, pol, Nx=3, Symmetry="none")
        f = O.rcwa_fields(dev, 0.94, 0.3, 0.4, pol, 0, 0, z, Nx=3)
        np.testing.assert_allclose(f["out"]["T"], o["T"], atol=1e-12)
        np.testing.assert_allclose(f["out"]["R"], o["R"], atol=1e-12)
        np.testing.assert_allclose(f["flux"], o["Ttot"][0] * np.ones(z.size), atol=1e-10)


def test_fields_tangential_components_are_continuous():
    """tests/TestRCWAFields.m:46-58."""
    dev = _field_pillars()
    x = np.arange(-0.2, 0.2001, 0.04)
    y = np.arange(-0.2, 0.2001, 0.1)
    dz = 1e-12
    f = O.rcwa_fields(dev, 0.94, 0.3, 0.4, "p", x, y, [-dz, dz, 0.5 - dz, 0.5 + dz], Nx=3)
    for c in range(2):
        for a, b in ((1, 0), (3, 2)):
            np.testing.assert_allclose(f["E"][:, :, a, c], f["E"][:, :, b, c], atol=1e-9)
            np.testing.assert_allclose(f["H"][:, :, a, c], f["H"][:, :, b, c], atol=1e-9)


def test_fields_lossy_layer_absorbs():
    """tests/TestRCWAFields.m:60-70: flux drops only inside the metal; 1 - R above, T below."""
    met = mc(0.2 + 3j)
    g = S.makeStructure([0.5, math.inf], mc(1), mc(1.45), [S.makeLayer(0.05, mc(1), [S.shapeRect(0.25, math.inf, met)])])
    f = O.rcwa_fields(g, 0.8, 0.2, 0, "p", 0, 0, [-0.1, 0, 0.025, 0.05, 0.3], Nx=15, Nsample=(1024, 1))
    assert abs(f["flux"][0] - (1 - f["out"]["Rtot"][0])) < 1e-10
    assert abs(f["flux"][-1] - f["out"]["Ttot"][0]) < 1e-10
    assert np.all(np.diff(f["flux"][1:4]) < 0)


def test_fields_plane_wave_transversality_and_impedance():
    """In a uniform medium the synthesised field is a plane wave: k . E = 0 and H~ = i k x E."""
    dev = S.makeStructure([0.5, 0.5], mc(1.3), mc(1.3), [S.makeLayer(0.2, mc(1.3))])
    th, ph = 0.4, 0.3
    f = O.rcwa_fields(dev, 0.8, th, ph, [0.6, 0.8j], 0.1, -0.05, [-0.2, 0.1, 0.5], Nx=1)
    kv = 1.3 * np.array([math.sin(th) * math.cos(ph), math.sin(th) * math.sin(ph), math.cos(th)])
    E = f["E"][0, 0]
    H = f["H"][0, 0]
    np.testing.assert_allclose(E @ kv, 0, atol=1e-12)
    np.testing.assert_allclose(H, 1j * np.cross(kv, E), atol=1e-12)
    np.testing.assert_allclose(np.linalg.norm(E, axis=1), 1.0, atol=1e-12)
