"""HDF5 measurement and density files in the reference's formats (CPU).

The density case restates test_H5DensityWriter.cpp (grid origin
(0.1, 0.2, 0.3), voxel (1, 2, 3), 5 x 6 x 7 voxels, one active voxel at
(1, 2, 3)) and checks every attribute the writer sets
(H5DensityWriter.hpp:97-231); the measurement case checks that the 5-D
"messung" index order is the channel id order of MeasurementSetupLinIndex and
that reading keeps exactly the non-zero channels (wrappers.hpp:329-368).
"""
import os

import numpy as np
import pytest

from raptr_amd import geometry as G

h5io = pytest.importorskip("raptr_amd.h5io")
if not os.path.exists(h5io._LIB):
    pytest.skip("HDF5 library not built", allow_module_level=True)


def test_density_writer_reference_case(tmp_path):
    grid = G.VoxelGrid((0.1, 0.2, 0.3), (1.0, 2.0, 3.0), (5, 6, 7))
    img = np.zeros((5, 6, 7), np.float32)
    img[1, 2, 3] = 1.0
    fn = str(tmp_path / "density.h5")
    h5io.write_density(fn, img, grid)
    assert h5io.dims(fn, "density") == (5, 6, 7)
    back, g = h5io.read_density(fn)
    np.testing.assert_array_equal(back, img)
    # the active voxel sits at the grid's linear index (z fastest)
    assert np.flatnonzero(back.reshape(-1)).tolist() == [grid.lin_id(1, 2, 3)]
    import ctypes as C
    lo = np.zeros(3, np.float32)
    hi = np.zeros(3, np.float32)
    n = np.zeros(3, np.int32)
    assert h5io.lib().raptr_h5_density_grid(fn.encode(), h5io._p(lo, C.c_float),
                                           h5io._p(hi, C.c_float), h5io._p(n, C.c_int)) == 0
    np.testing.assert_array_equal(lo, np.float32([0.1, 0.2, 0.3]))
    want_hi = np.float32([0.1, 0.2, 0.3]) + np.float32([5, 6, 7]) * np.float32([1, 2, 3])
    np.testing.assert_array_equal(hi, want_hi)
    assert n.tolist() == [5, 6, 7]
    assert g.n == (5, 6, 7)
    np.testing.assert_allclose(g.d, (1.0, 2.0, 3.0), rtol=1e-6)


def test_measurement_round_trip_and_channel_order(tmp_path):
    setup, _, _ = G.config_tiny()
    rng = np.random.default_rng(2)
    ch = np.sort(rng.choice(setup.nchannels, 200, replace=False)).astype(np.int32)
    counts = rng.integers(1, 50, len(ch)).astype(np.float32)
    fn = str(tmp_path / "meas.h5")
    h5io.write_measurement(fn, setup, ch, counts)
    assert h5io.dims(fn, "messung") == (setup.na, setup.n0z, setup.n0y, setup.n1z, setup.n1y)
    got_ch, got = h5io.read_measurement(fn, setup)
    np.testing.assert_array_equal(got_ch, ch)
    np.testing.assert_array_equal(got, counts)
    dense = h5io.read_measurement_dense(fn)
    for c in ch[:20]:
        ida, i0z, i0y, i1z, i1y = (int(v) for v in setup.split(c))
        assert dense[ida, i0z, i0y, i1z, i1y] == counts[np.searchsorted(ch, c)]
        assert setup.lin_id(i0z, i0y, i1z, i1y, ida) == c


def test_errors_are_reported(tmp_path):
    setup, grid, _ = G.config_tiny()
    fn = str(tmp_path / "density.h5")
    h5io.write_density(fn, np.zeros(grid.nvox, np.float32), grid)
    with pytest.raises(h5io.H5Error, match="messung"):
        h5io.read_measurement(fn)
    with pytest.raises(h5io.H5Error):
        h5io.read_measurement(str(tmp_path / "missing.h5"))
    mfn = str(tmp_path / "meas.h5")
    h5io.write_measurement(mfn, setup, [0], [1.0])
    other, _, _ = G.config_small()
    with pytest.raises(h5io.H5Error, match="setup needs"):
        h5io.read_measurement(mfn, other)
    with pytest.raises(h5io.H5Error):
        h5io.write_density(fn, np.zeros(3, np.float32), grid)


def test_cpp_mirror_client(tmp_path):
    """include/raptr/io.hpp compiled into tests/cpp/test_io.cpp (built by the
    package build) reads and writes the same files."""
    import subprocess
    exe = os.path.join(os.path.dirname(h5io._LIB), "test_io")
    if not os.path.exists(exe):
        pytest.skip("C++ io client not built")
    r = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().endswith("OK")


# ------------------------------------- against the reference's own HDF5 code
def _grid_attrs(fn):
    import ctypes as C
    lo, hi, n = np.zeros(3, np.float32), np.zeros(3, np.float32), np.zeros(3, np.int32)
    assert h5io.lib().raptr_h5_density_grid(fn.encode(), h5io._p(lo, C.c_float),
                                           h5io._p(hi, C.c_float), h5io._p(n, C.c_int)) == 0
    return lo, hi, n


def test_reads_density_written_by_reference():
    """tests/golden/density_ref.h5 was written by the reference's
    H5DensityWriter (compiled from its sources against HDF5 C++): values,
    shape and every grid attribute read back exactly."""
    from conftest import GOLDEN, load_golden
    gold = load_golden("h5_ref.npz")
    fn = os.path.join(GOLDEN, "density_ref.h5")
    nv = tuple(int(v) for v in gold["nvox"])
    assert h5io.dims(fn, "density") == nv
    back, g = h5io.read_density(fn)
    np.testing.assert_array_equal(back.reshape(-1), gold["density"])
    lo, hi, n = _grid_attrs(fn)
    np.testing.assert_array_equal(lo, gold["origin"])
    np.testing.assert_array_equal(hi, gold["origin"] + gold["nvox"].astype(np.float32) *
                                  gold["voxel"])
    assert n.tolist() == list(nv)
    # our writer gives a file with the same contents and attributes
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        mine = os.path.join(d, "d.h5")
        h5io.write_density(mine, gold["density"].reshape(nv),
                           G.VoxelGrid(tuple(gold["origin"]), tuple(gold["voxel"]), nv))
        for a, b in zip(_grid_attrs(mine), _grid_attrs(fn)):
            np.testing.assert_array_equal(a, b)
        np.testing.assert_array_equal(h5io.read_density(mine)[0], back)


def test_reference_reader_reads_our_measurement(oracle):
    """The measurement file raptr_amd.h5io writes is read by the reference's
    H5Reader ("messung", sizeOfFile, read) as the dense channel vector:
    recorded when the fixture was made, and live where the reference
    binding is built."""
    from conftest import load_golden
    import tempfile
    gold = load_golden("h5_ref.npz")
    setup = G.config_tiny()[0]
    dense = np.zeros(setup.nchannels, np.float32)
    dense[gold["meas_channels"]] = gold["meas_counts"]
    np.testing.assert_array_equal(gold["meas_ref_read"], dense)
    with tempfile.TemporaryDirectory() as d:
        fn = os.path.join(d, "m.h5")
        h5io.write_measurement(fn, setup, gold["meas_channels"], gold["meas_counts"])
        np.testing.assert_array_equal(h5io.read_measurement_dense(fn).reshape(-1), dense)
        if oracle.h5ref_available():
            np.testing.assert_array_equal(oracle.ref_h5_read(fn, False, setup.nchannels), dense)
            # and a density written here is read by the reference's
            # H5DensityReader
            grid = G.VoxelGrid((0.1, 0.2, 0.3), (1.0, 2.0, 3.0), (5, 6, 7))
            img = np.arange(210, dtype=np.float32).reshape(5, 6, 7) * np.float32(0.5)
            dfn = os.path.join(d, "d.h5")
            h5io.write_density(dfn, img, grid)
            np.testing.assert_array_equal(oracle.ref_h5_read(dfn, True, 210), img.reshape(-1))
