"""The reference's HDF5 file formats (host side, libraptr_h5.so).

* Measurement files: dataset "messung", float32, 5-D (angle, det0 z,
  det0 y, det1 z, det1 y); the row-major linear index is the channel id.
  ``read_measurement`` returns the non-zero channels and their counts, as
  readHDF5_MeasVct / H5File2DefaultMeasurementList do (wrappers.hpp:329-368,
  H5File2DefaultMeasurementList.h).
* Density files: dataset "density", float32, (xnbin, ynbin, znbin) with
  the grid attributes of H5DensityWriter.hpp:97-231.
"""
import ctypes as C
import os

import numpy as np

from .geometry import VoxelGrid

_LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libraptr_h5.so")
_lib = None


class H5Error(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            raise H5Error(f"HDF5 support not built: {_LIB} missing (needs the HDF5 C library)")
        L = C.CDLL(_LIB)
        P, i32, i64, f32 = C.POINTER, C.c_int, C.c_int64, C.c_float
        sig = {
            "raptr_h5_last_error": (C.c_char_p, []),
            "raptr_h5_dims": (i32, [C.c_char_p, C.c_char_p, P(i32), P(i64), i32]),
            "raptr_h5_read": (i32, [C.c_char_p, C.c_char_p, P(f32), i64]),
            "raptr_h5_write_measurement": (i32, [C.c_char_p, P(f32), P(i32)]),
            "raptr_h5_write_density": (i32, [C.c_char_p, P(f32), P(f32), P(f32), P(i32)]),
            "raptr_h5_density_grid": (i32, [C.c_char_p, P(f32), P(f32), P(i32)]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype, fn.argtypes = res, args
        _lib = L
    return _lib


def _check(rc, what):
    if rc != 0:
        raise H5Error(f"{what}: {lib().raptr_h5_last_error().decode()}")


def _p(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def dims(path, dataset):
    nd = C.c_int()
    d = np.zeros(8, np.int64)
    _check(lib().raptr_h5_dims(path.encode(), dataset.encode(), C.byref(nd), _p(d, C.c_int64), 8),
           "dims")
    return tuple(int(v) for v in d[:nd.value])


def read_measurement_dense(path, setup=None):
    """The full "messung" array (float32, 5-D).  With a setup, its extents
    must be (na, n0z, n0y, n1z, n1y)."""
    shape = dims(path, "messung")
    if len(shape) != 5:
        raise H5Error(f"{path}: messung has {len(shape)} dimensions, not 5")
    if setup is not None:
        want = (setup.na, setup.n0z, setup.n0y, setup.n1z, setup.n1y)
        if shape != want:
            raise H5Error(f"{path}: messung is {shape}, the setup needs {want}")
    out = np.empty(shape, np.float32)
    _check(lib().raptr_h5_read(path.encode(), b"messung", _p(out, C.c_float), out.size),
           "read messung")
    return out


def read_measurement(path, setup=None):
    """(channel ids int32 ascending, counts float32) of the channels with a
    non-zero count."""
    dense = read_measurement_dense(path, setup).reshape(-1)
    ch = np.nonzero(dense != 0)[0].astype(np.int32)
    return ch, dense[ch].copy()


def write_measurement(path, setup, channels, counts):
    """Write counts of the given channels (others 0) as a "messung" file."""
    dense = np.zeros(setup.nchannels, np.float32)
    dense[np.asarray(channels, np.int64)] = np.asarray(counts, np.float32)
    d = np.array([setup.na, setup.n0z, setup.n0y, setup.n1z, setup.n1y], np.int32)
    _check(lib().raptr_h5_write_measurement(path.encode(), _p(dense, C.c_float),
                                            _p(d, C.c_int)), "write messung")


def write_density(path, image, grid):
    img = np.ascontiguousarray(image, np.float32).reshape(-1)
    if img.size != grid.nvox:
        raise H5Error(f"image has {img.size} voxels, grid {grid.nvox}")
    o = np.array(grid.o, np.float32)
    d = np.array(grid.d, np.float32)
    n = np.array(grid.n, np.int32)
    _check(lib().raptr_h5_write_density(path.encode(), _p(img, C.c_float), _p(o, C.c_float),
                                        _p(d, C.c_float), _p(n, C.c_int)), "write density")


def read_density(path):
    """(image float32 of shape (xnbin, ynbin, znbin), VoxelGrid from the
    attributes: origin = min, pitch = (max - min) / nbin)."""
    lo = np.zeros(3, np.float32)
    hi = np.zeros(3, np.float32)
    n = np.zeros(3, np.int32)
    _check(lib().raptr_h5_density_grid(path.encode(), _p(lo, C.c_float), _p(hi, C.c_float),
                                       _p(n, C.c_int)), "density grid")
    img = np.empty(tuple(int(v) for v in n), np.float32)
    _check(lib().raptr_h5_read(path.encode(), b"density", _p(img, C.c_float), img.size),
           "read density")
    pitch = tuple(float((hi[k] - lo[k]) / n[k]) for k in range(3))
    return img, VoxelGrid(tuple(float(v) for v in lo), pitch, tuple(int(v) for v in n))
