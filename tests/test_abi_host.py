"""C-ABI checks that need no GPU: the library loads, exports every function
include/hamsx.h declares, rejects bad arguments with the documented codes,
and its host-side set-up routines (panel geometry, wave properties) agree
bit for bit with the oracle."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import oracle
from conftest import ROOT
from hamsx import _lib
from hamsx.mesh import box, combine, cylinder

HEADER = os.path.join(ROOT, "include", "hamsx.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hamsx_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(_lib.SIGNATURES), set(names) ^ set(_lib.SIGNATURES)


def test_version_and_argument_errors():
    lib = _lib.load()
    assert lib.hamsx_version().startswith(b"hamsx")
    assert lib.hamsx_device_count() >= 0
    # a NULL context is an argument error, never a crash
    assert lib.hamsx_destroy(None) in (_lib.HAMSX_OK, _lib.HAMSX_EARG)
    assert lib.hamsx_assemble(None) == _lib.HAMSX_EARG
    assert lib.hamsx_set_precision(None, 0) == _lib.HAMSX_EARG
    assert lib.hamsx_create(None, 0) == _lib.HAMSX_EARG
    v, wk, w1 = C.c_double(), C.c_double(), C.c_double()
    wvn = np.zeros(_lib.HAMSX_NK)
    assert lib.hamsx_wave_properties(9, 1.0, 50.0, C.byref(v), C.byref(wk), C.byref(w1),
                                     _lib.ptr(wvn)) == _lib.HAMSX_EARG
    assert b"1..5" in lib.hamsx_last_error()


@pytest.mark.parametrize("w,h", [(0.2, 50.0), (1.0, -1.0), (2.5, 50.0), (0.1, 200.0), (3.0, 30.0)])
def test_wave_properties_bitwise_equal_oracle(w, h):
    from hamsx.solver import wave_properties
    v, wk, w1, wvn = wave_properties(3, w, h)
    vo, wko, wvno = oracle.wave(w, h)
    assert v == vo and wk == wko and w1 == w
    np.testing.assert_array_equal(wvn, wvno)


@pytest.mark.parametrize("body", ["cylinder", "box"])
def test_panel_geometry_bitwise_equal_oracle(body):
    b = cylinder(4.0, 7.0, 20, 5, 3, 2.0, -1.0) if body == "cylinder" else box(12.0, 8.0, 5.0, 6, 4, 3)
    m = combine([b], [np.array([1.0, -0.5, -2.0])])
    cen, area, pnsz, dxyz = oracle.geometry(b, [1.0, -0.5, -2.0])
    np.testing.assert_array_equal(m.cen, cen)
    np.testing.assert_array_equal(m.area, area)
    np.testing.assert_array_equal(m.pnsz, pnsz)
    np.testing.assert_array_equal(m.dxyz, dxyz)
