"""GPU tests of the HDF5-file drivers (reco.cu, backprojection.cu
restated over the device path): their images equal the library calls they
wrap, per-iteration files are written as the reference names them."""
import os

import numpy as np
import pytest

from raptr_amd import geometry as G

pytestmark = pytest.mark.gpu
h5io = pytest.importorskip("raptr_amd.h5io")


@pytest.fixture(scope="module")
def tiny_files(gpu, tmp_path_factory):
    if not os.path.exists(h5io._LIB):
        pytest.skip("HDF5 library not built")
    from raptr_amd import synthetic
    setup, grid, nrays = G.config_tiny()
    d = tmp_path_factory.mktemp("drv")
    img = synthetic.spheres_phantom(grid)
    meas, y, _ = synthetic.measurement(setup, grid, nrays, img, 1e5, seed=3)
    fn = str(d / "meas.h5")
    h5io.write_measurement(fn, setup, meas, y)
    return d, fn, setup, grid, nrays, meas, y


def test_reco_driver_equals_library(tiny_files):
    from raptr_amd import drivers, sysmat
    from raptr_amd.recon import Reconstructor
    d, fn, setup, grid, nrays, meas, y = tiny_files
    out = str(d / "out.h5")
    x = drivers.reco(fn, out, setup, grid, nrays, 4, every_iteration=False)
    A = sysmat.build_csr(setup, grid, meas, nrays)
    R = Reconstructor(A, sysmat.sensitivity(setup, grid, nrays), "f32", 1, row_order=setup)
    want = R.mlem(y, np.full(grid.nvox, 1.0 / grid.nvox, np.float32), 4, use_graph=False)
    np.testing.assert_array_equal(x, want)
    back, g = h5io.read_density(out)
    np.testing.assert_array_equal(back.reshape(-1), want)
    assert g.n == grid.n


def test_reco_driver_writes_every_iteration(tiny_files):
    from raptr_amd import drivers
    d, fn, setup, grid, nrays, _, _ = tiny_files
    out = str(d / "it.h5")
    x = drivers.reco(fn, out, setup, grid, nrays, 3)
    for it in range(3):
        assert os.path.exists(str(d / f"{it}_it.h5"))
    last, _ = h5io.read_density(str(d / "2_it.h5"))
    np.testing.assert_array_equal(last.reshape(-1), x)
    assert abs(float(x.astype(np.float64).sum()) - 1.0) < 1e-5


def test_reco_driver_sensitivity_and_guess_files(tiny_files):
    from raptr_amd import drivers, sysmat
    d, fn, setup, grid, nrays, _, _ = tiny_files
    s = sysmat.sensitivity(setup, grid, nrays).numpy()
    sfn, gfn = str(d / "sens.h5"), str(d / "guess.h5")
    h5io.write_density(sfn, s.astype(np.float32), grid)
    h5io.write_density(gfn, np.full(grid.nvox, 1.0 / grid.nvox, np.float32), grid)
    a = drivers.reco(fn, str(d / "a.h5"), setup, grid, nrays, 4, every_iteration=False)
    b = drivers.reco(fn, str(d / "b.h5"), setup, grid, nrays, 4, sensitivity=sfn, guess=gfn,
                     every_iteration=False)
    assert np.linalg.norm(a - b) / np.linalg.norm(a) < 1e-5


def test_backprojection_driver(tiny_files):
    from raptr_amd import drivers, sysmat
    d, fn, setup, grid, nrays, meas, y = tiny_files
    out = str(d / "bp.h5")
    x = drivers.backprojection(fn, out, setup, grid, nrays)
    want = sysmat.backproject(setup, grid, nrays, meas, y)
    np.testing.assert_array_equal(x, want)
    back, _ = h5io.read_density(out)
    np.testing.assert_array_equal(back.reshape(-1), want.astype(np.float32))
