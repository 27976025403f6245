"""MLEM operations (mlemOperations.hpp:44-169) and the plain MLEM driver
(reco.cu:230-303) on a device-resident matrix."""
import ctypes as C

import numpy as np

from . import _capi as capi
from .device import DeviceArray, Workspace


def divides(a, b, stream=None):
    da, db = DeviceArray.from_numpy(a, np.float32), DeviceArray.from_numpy(b, np.float32)
    r = DeviceArray(len(da), np.float32)
    capi.call("raptr_divides", r.ptr, da.ptr, db.ptr, len(da), stream)
    return r.numpy()


def divides_multiplies(a, b, c, stream=None):
    da, db, dc = (DeviceArray.from_numpy(v, np.float32) for v in (a, b, c))
    r = DeviceArray(len(da), np.float32)
    capi.call("raptr_divides_multiplies", r.ptr, da.ptr, db.ptr, dc.ptr, len(da), stream)
    return r.numpy()


def scales(x, alpha, stream=None):
    dx = DeviceArray.from_numpy(x, np.float32)
    capi.call("raptr_scales", dx.ptr, float(alpha), len(dx), stream)
    return dx.numpy()


def fsum(x, stream=None):
    dx = DeviceArray.from_numpy(x, np.float32)
    ws = Workspace(capi.lib().raptr_sum_workspace_size(len(dx)))
    out = C.c_float()
    capi.call("raptr_sum", dx.ptr, len(dx), C.byref(out), ws.ptr, ws.nbytes, stream)
    return float(out.value)


def mlem(A, At, y, sens, x0, niter, use_graph=True, stream=None, cfg=(0, 0)):
    """Plain MLEM.  A, At: DeviceCSR of the measured rows and its transpose;
    y: counts of the measured rows (host or device float32); sens: FP64
    sensitivity (device); x0: start image (host).  Returns the image."""
    dy = y if isinstance(y, DeviceArray) else DeviceArray.from_numpy(y, np.float32)
    dx = DeviceArray.from_numpy(x0, np.float32)
    ws = Workspace(capi.lib().raptr_mlem_workspace_size(A.m, A.n))
    capi.call("raptr_mlem", A.m, A.n, A.row_ptr.ptr, A.col.ptr, A.val.ptr, At.row_ptr.ptr,
              At.col.ptr, At.val.ptr, dy.ptr, sens.ptr, dx.ptr, int(niter),
              1 if use_graph else 0, int(cfg[0]), int(cfg[1]), ws.ptr, ws.nbytes, stream)
    return dx.numpy()


def autotune(A, At, y, x, reps=5, stream=None):
    """Fastest (forward, backward) kernel configuration codes and their
    times in ms on this matrix."""
    dy = y if isinstance(y, DeviceArray) else DeviceArray.from_numpy(y, np.float32)
    dx = x if isinstance(x, DeviceArray) else DeviceArray.from_numpy(x, np.float32)
    ws = Workspace(capi.lib().raptr_mlem_workspace_size(A.m, A.n))
    f, b = C.c_int(), C.c_int()
    fm, bm = C.c_float(), C.c_float()
    capi.call("raptr_mlem_autotune", A.m, A.n, A.row_ptr.ptr, A.col.ptr, A.val.ptr,
              At.row_ptr.ptr, At.col.ptr, At.val.ptr, dy.ptr, dx.ptr, int(reps), C.byref(f),
              C.byref(b), C.byref(fm), C.byref(bm), ws.ptr, ws.nbytes, stream)
    return (f.value, b.value), (fm.value, bm.value)


def uniform_image(nvox):
    """x0 = 1/N (reco.cu:135-158 without a guess file)."""
    return np.full(nvox, 1.0 / nvox, np.float32)
