"""GPU field reconstruction (orf_fields through orf.rcwaFields) against the
FP64 oracle and the physics of tests/TestRCWAFields.m."""
import math

import numpy as np
import pytest

import opticrcwafdtd_amd as orf
from oracle import rcwa_oracle as O

pytestmark = pytest.mark.gpu
mc = orf.materialConst


def _pillars():
    return orf.makeStructure([0.4, 0.4], mc(1.45), mc(1),
                             [orf.makeLayer(0.5, mc(1), [orf.shapeCircle(0.1, mc(3.6))]), orf.makeLayer(0.2, mc(2))])


@pytest.mark.parametrize("pol", ["s", "p", [0.6, 0.8j]])
def test_fields_match_oracle(pol):
    dev = _pillars()
    x = np.linspace(-0.2, 0.2, 7)
    y = np.linspace(-0.2, 0.15, 5)
    z = np.array([-0.3, -0.01, 0.0, 0.1, 0.25, 0.499, 0.5, 0.6, 0.7, 0.9])
    g = orf.rcwaFields(dev, 0.94, 0.3, 0.4, pol, x, y, z, Nx=3, Nsample=(128, 128))
    o = O.rcwa_fields(dev, 0.94, 0.3, 0.4, pol, x, y, z, Nx=3, Nsample=(128, 128))
    assert g["E"].shape == (7, 5, 10, 3)
    np.testing.assert_allclose(g["E"], o["E"], atol=1e-10, rtol=0)
    np.testing.assert_allclose(g["H"], o["H"], atol=1e-10, rtol=0)
    np.testing.assert_allclose(g["flux"], o["flux"], atol=1e-10, rtol=0)
    np.testing.assert_allclose(g["eps"], o["eps"], atol=0, rtol=0)
    np.testing.assert_allclose(g["out"]["T"][:, 0], o["out"]["T"][0], atol=1e-10)
    np.testing.assert_allclose(g["flux"], g["out"]["Ttot"][0], atol=1e-10)  # lossless: flux = Ttot everywhere


def test_fields_slab_analytic_and_lossy_absorption():
    nidx = [1, 2, 1.45]
    d, k = 0.3, 2 * np.pi
    slab = orf.makeStructure([0.37, 0.37], mc(nidx[0]), mc(nidx[2]), [orf.makeLayer(d, mc(nidx[1]))])
    z = np.linspace(-0.4, 0.7, 23)
    f = orf.rcwaFields(slab, 1, 0, 0, "s", 0, 0, z, Nx=1)
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
    met = mc(0.2 + 3j)
    g = orf.makeStructure([0.5, math.inf], mc(1), mc(1.45),
                          [orf.makeLayer(0.05, mc(1), [orf.shapeRect(0.25, math.inf, met)])])
    f = orf.rcwaFields(g, 0.8, 0.2, 0, "p", 0, 0, [-0.1, 0, 0.025, 0.05, 0.3], Nx=15, Nsample=(1024, 1))
    assert abs(f["flux"][0] - (1 - f["out"]["Rtot"][0])) < 1e-10
    assert abs(f["flux"][-1] - f["out"]["Ttot"][0]) < 1e-10
    assert np.all(np.diff(f["flux"][1:4]) < 0)


@pytest.mark.parametrize("sym", ["none", "auto"])
def test_single_layer_direct_solve_equals_smatrix_route(sym):
    """One patterned layer is solved on its mode amplitudes without a layer
    S-matrix; the kept-modes plan behind rcwaFields forms the layer S-matrix and
    star products (smatrixLayer.m:6-18, smatrixStar.m:4-10).  Same R and T."""
    dev = orf.makeStructure([0.4, 0.4], mc(1.45), mc(1),
                            [orf.makeLayer(0.6, mc(1), [orf.shapeCircle(0.12, mc(3.6 + 0.01j))])])
    for pol in ("s", "p"):
        g = orf.rcwa2D(dev, 0.94, 0, 0, pol, Nx=5, Nsample=(256, 256), Symmetry=sym)
        f = orf.rcwaFields(dev, 0.94, 0, 0, pol, np.zeros(1), np.zeros(1), np.array([0.3]), Nx=5,
                           Nsample=(256, 256))["out"]
        assert np.array_equal(np.ravel(g["m"]), np.ravel(f["m"])) and np.array_equal(np.ravel(g["n"]), np.ravel(f["n"]))
        dT = np.max(np.abs(np.ravel(g["T"]) - np.ravel(f["T"])))
        dR = np.max(np.abs(np.ravel(g["R"]) - np.ravel(f["R"])))
        print(f"direct vs S-matrix route ({sym}, {pol}): max |dT| {dT:.2e}, max |dR| {dR:.2e}")
        assert dT < 1e-11 and dR < 1e-11
