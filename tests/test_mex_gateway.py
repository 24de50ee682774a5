"""The MATLAB MEX gateway (opticrcwafdtd_amd/matlab/orf_rcwa_mex.c) compiled
against a stand-in MEX API (tests/mex_shim/mex.h, fake_mex.c) and driven by a
small C program (tests/mex_shim/mex_driver.c).

CPU: linked against a stand-in library that returns a chosen status, so the
status -> MATLAB identifier mapping of include/orf_rcwa.h, the rethrow of the
progress function's own exception (tests/TestGUISupport.m:15-18) and the
argument checks run without a GPU.  GPU: linked against liborf_rcwa.so, a
solve through the gateway equals the Python path, and a cancel is rethrown.
"""
import json
import math
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SHIM = os.path.join(ROOT, "tests", "mex_shim")
GATEWAY = os.path.join(ROOT, "opticrcwafdtd_amd", "matlab", "orf_rcwa_mex.c")
LIBDIR = os.path.join(ROOT, "opticrcwafdtd_amd")

pytestmark = pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")


def _build(tmp, real_library):
    inc = ["-I" + SHIM, "-I" + os.path.join(ROOT, "include")]
    exe = os.path.join(tmp, "drv")
    srcs = [os.path.join(SHIM, "mex_driver.c"), os.path.join(SHIM, "fake_mex.c"), GATEWAY]
    if real_library:
        link = ["-L" + LIBDIR, "-lorf_rcwa", "-Wl,-rpath," + LIBDIR]
    else:
        subprocess.run(["gcc", "-O1", "-shared", "-fPIC", *inc, "-o", os.path.join(tmp, "liborf_stub.so"),
                        os.path.join(SHIM, "orf_stub.c")], check=True)
        link = ["-L" + tmp, "-lorf_stub", "-Wl,-rpath," + tmp]
    subprocess.run(["gcc", "-O1", "-Wall", *inc, "-o", exe, *srcs, *link, "-lm"], check=True,
                   capture_output=True)
    return exe


def _run(exe, mode, status=None, timeout=60):
    env = dict(os.environ)
    if status is not None:
        env["ORF_STUB_STATUS"] = str(status)
    p = subprocess.run([exe, mode], capture_output=True, text=True, env=env, timeout=timeout)
    assert p.returncode == 0, p.stderr
    return json.loads(p.stdout.strip().splitlines()[-1]), p.stderr


@pytest.fixture(scope="module")
def stub_driver(tmp_path_factory):
    return _build(str(tmp_path_factory.mktemp("mex_stub")), real_library=False)


def test_status_codes_map_to_reference_identifiers(stub_driver):
    ids = {1: "orf:invalidArgument", 2: "rcwa:pol", 3: "rcwa:lossyIncident", 4: "orf:outOfMemory",
           5: "orf:cancelled", 6: "orf:device", 7: "convMatrix1D:sampling"}
    for code, ident in ids.items():
        out, _ = _run(stub_driver, "solve", status=code)
        assert out["status"] == "error" and out["error_id"] == ident, (code, out)
        assert out["message"] == f"stub status {code}" and not out["rethrown"]


def test_solve_fills_the_output_struct(stub_driver):
    out, _ = _run(stub_driver, "solve", status=0)
    assert out["status"] == "ok" and out["M"] == 25
    assert out["R"] == [0.25] * 50 and out["persistent"] == 0


def test_progress_exception_is_rethrown_without_leaking(stub_driver):
    """A progress function that throws cancels the solve; its own MException (here
    'user:stop') is rethrown, and no persistent array is left behind."""
    out, _ = _run(stub_driver, "cancel")
    assert out == {"status": "error", "error_id": "user:stop", "message": "rethrown from the progress function",
                   "rethrown": True, "persistent": 0}
    out, err = _run(stub_driver, "progress")
    assert out["status"] == "ok"
    assert [float(l.split()[1]) for l in err.splitlines() if l.startswith("progress")] == [0.5, 1.0]


def test_missing_problem_field_is_an_invalid_argument(stub_driver):
    out, _ = _run(stub_driver, "missing")
    assert out["error_id"] == "orf:invalidArgument" and "'Lx'" in out["message"]


@pytest.mark.gpu
def test_gateway_on_the_device_equals_python_path(tmp_path):
    import opticrcwafdtd_amd as orf
    exe = _build(str(tmp_path), real_library=True)
    out, _ = _run(exe, "solve")
    assert out["status"] == "ok" and out["M"] == 25
    mc = orf.materialConst
    devs = [orf.makeStructure([0.4, 0.4], mc(1.45), mc(1),
                              [orf.makeLayer(0.6, mc(1), [orf.shapeCircle(r, mc(3.6 + 0.01j))])]) for r in (0.06, 0.12)]
    g = orf.rcwa2D(devs, 0.94, 0, 0, "s", Nx=2, Nsample=(64, 64), Symmetry="none")
    R = np.array(out["R"]).reshape(2, 25).T          # M x B, column-major
    T = np.array(out["T"]).reshape(2, 25).T
    t = np.array(out["t"]).view(complex).reshape(2, 25, 3).transpose(2, 1, 0)  # 3 x M x B
    np.testing.assert_allclose(R, g["R"], atol=1e-13, rtol=0)
    np.testing.assert_allclose(T, g["T"], atol=1e-13, rtol=0)
    np.testing.assert_allclose(t, g["t"], atol=1e-13, rtol=0)
    assert math.isclose(float(np.sum(R + T)), float(np.sum(g["Rtot"] + g["Ttot"])), abs_tol=1e-12)
    out, _ = _run(exe, "cancel")
    assert out["error_id"] == "user:stop" and out["rethrown"] and out["persistent"] == 0
    out, _ = _run(exe, "badpol")
    assert out["error_id"] == "rcwa:pol"
