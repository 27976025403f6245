"""CPU checks of the C-ABI libraries: each loads and exports every function
its header declares (include/raptr_hip.h -> libraptr_hip.so,
include/raptr_h5.h -> libraptr_h5.so), and the Python declarations cover
them."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions(header="raptr_hip.h"):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(raptr_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from raptr_amd import _capi
    L = C.CDLL(_capi.LIB_PATH)
    names = declared_functions()
    assert len(names) > 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_h5_library_exports_every_declared_symbol():
    from raptr_amd import h5io
    if not os.path.exists(h5io._LIB):
        pytest.skip("HDF5 library not built")
    L = C.CDLL(h5io._LIB)
    names = declared_functions("raptr_h5.h")
    assert len(names) >= 6
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_python_bindings_cover_header():
    from raptr_amd import _capi
    assert set(declared_functions()) <= set(_capi.exported_symbols())


def test_status_strings_without_device():
    from raptr_amd import _capi
    L = _capi.lib()
    assert L.raptr_status_string(0) == b"ok"
    assert L.raptr_status_string(_capi.RAPTR_ERR_WORKSPACE) == b"workspace too small"
    assert L.raptr_version() >= 100


def test_host_geometry_helpers_match_oracle(oracle):
    """The library's host-side angle table and cylinder radius equal the
    oracle's (no device needed)."""
    import numpy as np
    from raptr_amd import _capi, geometry as G
    L = _capi.lib()
    for setup, grid, _ in (G.config_small(), G.config_full()):
        s = setup.struct()
        sn = np.zeros(setup.na, np.float32)
        cs = np.zeros(setup.na, np.float32)
        assert L.raptr_angle_table(C.byref(s), sn.ctypes.data_as(C.POINTER(C.c_float)),
                                   cs.ctypes.data_as(C.POINTER(C.c_float))) == 0
        os_ = oracle.make_setup(setup.pos0x, setup.pos1x, setup.na, setup.n0z, setup.n0y,
                                setup.n1z, setup.n1y, setup.da, setup.segx, setup.segy,
                                setup.segz)
        osn, ocs = oracle.angle_table(os_)
        np.testing.assert_array_equal(sn, osn)
        np.testing.assert_array_equal(cs, ocs)
        g = grid.struct()
        r = L.raptr_sum_radii(C.byref(s), C.byref(g))
        want = np.float32(0.5) * (np.float32(np.sqrt(np.float32(3) * np.float32(grid.d[0]) ** 2))
                                  + np.float32(np.sqrt(np.float32(setup.segx ** 2 + setup.segy ** 2
                                                                  + setup.segz ** 2))))
        assert r == pytest.approx(float(want), rel=1e-6)


def test_invalid_geometry_is_rejected_without_device():
    """Argument validation runs before any device call."""
    from raptr_amd import _capi, geometry as G
    L = _capi.lib()
    setup, grid, _ = G.config_tiny()
    bad = G.MeasurementSetup(-230, 230, 0, 4, 4, 4, 4, 60, 20, 4, 4)
    s, g = bad.struct(), grid.struct()
    st = L.raptr_sm_from_wq(C.byref(s), C.byref(g), None, None, None, 10, 4, 1, 0, None)
    assert st == _capi.RAPTR_ERR_INVALID_ARG
    assert b"na=0" in L.raptr_last_error()
    s = setup.struct()
    st = L.raptr_sm_from_wq(C.byref(s), C.byref(g), None, None, None, 10, 0, 1, 0, None)
    assert st == _capi.RAPTR_ERR_INVALID_ARG
