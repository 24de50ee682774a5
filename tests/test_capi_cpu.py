"""CPU-side checks of the C-ABI library: it loads, exports every entry point
the headers declare, and the host-side flattening matches the header layout.
No GPU is touched."""
import math
import os
import re

import numpy as np
import pytest

from opticrcwafdtd_amd import _lib
import opticrcwafdtd_amd as orf

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in ("orf_rcwa.h", "orf_pageops.h"):
        src = open(os.path.join(ROOT, "include", h)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        names |= set(re.findall(r"\b(orf_[a-z0-9_]+)\s*\(", src))
    names.discard("orf_progress_fn")
    return names


def test_library_exports_every_declared_symbol():
    lib = _lib.load_library()
    declared = _declared()
    assert declared, "no declarations found"
    missing = [n for n in sorted(declared) if not hasattr(lib, n)]
    assert not missing, missing
    assert declared <= set(_lib.EXPORTS) | {"orf_progress_fn"}


def test_harmonic_count_and_flattening():
    lib = _lib.load_library()
    mc = orf.materialConst
    devs = [orf.makeStructure([0.4, 0.4], mc(1.45), mc(1),
                              [orf.makeLayer(0.6, mc(1), [orf.shapeCircle(r, mc(3.6 + 0.01j))])])
            for r in (0.05, 0.1)]
    p = orf.Problem(devs, [0.9, 0.94], 0, 0, ["s", "p"], Nx=3)
    assert lib.orf_harmonic_count(p.struct) == 49
    assert p.struct.nLambda == 2 and p.struct.nPol == 2 and p.struct.nStruct == 2
    assert p.struct.nMaterials == 3   # constant materials are deduplicated by value
    eps = p._eps.reshape(2, 3)        # [lambda][material]
    assert abs(eps[0, 0] - 1.45 ** 2) < 1e-15
    g = orf.makeStructure([0.7, math.inf], mc(1), mc(1.45), [orf.makeLayer(0.5, mc(1))])
    q = orf.Problem(g, 1.0, 0.1, 0, "p", Nx=20)
    assert q.struct.Ny == 0 and q.struct.Nsy == 1 and lib.orf_harmonic_count(q.struct) == 41


def test_invalid_pol_rejected_on_host():
    mc = orf.materialConst
    g = orf.makeStructure([0.5, 0.5], mc(1), mc(1.45), [orf.makeLayer(0.5, mc(1))])
    with pytest.raises(orf.OrfError) as e:
        orf.Problem(g, 1.0, 0, 0, "x", Nx=1)
    assert e.value.identifier == "rcwa:pol"


def test_tuning_table_loads_and_rejects_bad_lines(tmp_path):
    lib = _lib.load_library()
    good = tmp_path / "t.txt"
    good.write_text("# comment\n0 61 0 1\n1 242 1 12   # trailing note\n2 98 0 16\n")
    assert lib.orf_tune_load(str(good).encode()) == _lib.ORF_OK
    bad = tmp_path / "b.txt"
    bad.write_text("0 61 0 1\nnonsense\n")
    assert lib.orf_tune_load(str(bad).encode()) == _lib.ORF_INVALID_ARG
    assert "line 2" in _lib.last_error()
    assert lib.orf_tune_set(7, 10, 0, 1) == _lib.ORF_INVALID_ARG
    lib.orf_tune_clear()
    if os.path.exists(_lib.TUNING_TABLE):   # the committed table parses
        assert lib.orf_tune_load(_lib.TUNING_TABLE.encode()) == _lib.ORF_OK
        lib.orf_tune_clear()
