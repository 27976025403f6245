"""Workqueue and system-matrix construction on the GPU (Python side of the C
ABI).  Mirrors getWorkqueue (getWorkqueue.hpp:135-243), the
getSystemMatrixFromWorkqueue kernel (getSystemMatrixFromWorkqueue.cu:
116-281) and systemMatrixCalculation (wrappers.hpp:185-266)."""
import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _capi as capi
from .device import DeviceArray, Workspace


def workqueue(setup, grid, cnl_list, list_id=0, vxl_id=0, n=-1, stream=None):
    """Resumable (optionally bounded) workqueue.  Returns host arrays
    (cnl, vxl) and the advanced cursor (list_id, vxl_id)."""
    lst = np.ascontiguousarray(cnl_list, np.int32)
    d_list = DeviceArray.from_numpy(lst)
    s, g = setup.struct(), grid.struct()
    L = capi.lib()
    ws = Workspace(L.raptr_workqueue_workspace_size(C.byref(g), len(lst)))
    if n >= 0:
        cap = int(n)
    else:
        cnt = workqueue_counts(setup, grid, lst[list_id:] if list_id < len(lst) else lst[:0],
                               d_list=None)
        cap = int(cnt.sum())
    d_cnl = DeviceArray(max(cap, 1), np.int32)
    d_vxl = DeviceArray(max(cap, 1), np.int32)
    li, vi, found = C.c_int(list_id), C.c_int(vxl_id), C.c_int64()
    capi.call("raptr_workqueue", C.byref(s), C.byref(g), d_list.ptr, len(lst), C.byref(li),
              C.byref(vi), int(n), d_cnl.ptr, d_vxl.ptr, cap, C.byref(found), ws.ptr,
              ws.nbytes, stream)
    k = found.value
    return d_cnl.numpy(k), d_vxl.numpy(k), li.value, vi.value


def workqueue_counts(setup, grid, cnl_list, d_list=None, stream=None):
    lst = np.ascontiguousarray(cnl_list, np.int32)
    if len(lst) == 0:
        return np.zeros(0, np.int64)
    d_list = d_list or DeviceArray.from_numpy(lst)
    out = DeviceArray(len(lst), np.int64)
    s, g = setup.struct(), grid.struct()
    capi.call("raptr_workqueue_counts", C.byref(s), C.byref(g), d_list.ptr, len(lst), out.ptr,
              stream)
    return out.numpy()


def sm_from_wq(setup, grid, cnl, vxl, nrays, seed=1, raygen=capi.RAYGEN_RAND, stream=None):
    cnl = np.ascontiguousarray(cnl, np.int32)
    vxl = np.ascontiguousarray(vxl, np.int32)
    d_c, d_v = DeviceArray.from_numpy(cnl), DeviceArray.from_numpy(vxl)
    d_val = DeviceArray(max(len(cnl), 1), np.float32)
    s, g = setup.struct(), grid.struct()
    capi.call("raptr_sm_from_wq", C.byref(s), C.byref(g), d_c.ptr, d_v.ptr, d_val.ptr,
              len(cnl), int(nrays), int(seed), int(raygen), stream)
    return d_val.numpy(len(cnl))


def generate_rays(setup, cnl, nrays, seed=1, raygen=capi.RAYGEN_RAND, stream=None):
    cnl = np.ascontiguousarray(cnl, np.int32)
    d_c = DeviceArray.from_numpy(cnl)
    d_r = DeviceArray(max(len(cnl) * nrays * 6, 1), np.float32)
    s = setup.struct()
    capi.call("raptr_generate_rays", C.byref(s), d_c.ptr, len(cnl), int(nrays), int(seed),
              int(raygen), d_r.ptr, stream)
    return d_r.numpy(len(cnl) * nrays * 6).reshape(len(cnl), nrays, 6)


def get_chords(rays, grid, cap=None, stream=None):
    """Siddon chords of rays [k, 6] (getChords, ChordsCalc_lowlevel.hpp:
    194-379): returns (chords [k, cap], voxel ids [k, cap, 3], counts [k])."""
    rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 6)
    k = len(rays)
    cap = int(cap if cap is not None else sum(grid.n) + 4)
    d_r = DeviceArray.from_numpy(rays.reshape(-1) if k else np.zeros(6, np.float32))
    d_ch = DeviceArray(max(k * cap, 1), np.float32)
    d_id = DeviceArray(max(k * cap * 3, 1), np.int32)
    d_n = DeviceArray(max(k, 1), np.int32)
    g = grid.struct()
    capi.call("raptr_get_chords", d_r.ptr, k, C.byref(g), cap, d_ch.ptr, d_id.ptr, d_n.ptr,
              stream)
    if k == 0:
        return (np.zeros((0, cap), np.float32), np.zeros((0, cap, 3), np.int32),
                np.zeros(0, np.int32))
    return (d_ch.numpy(k * cap).reshape(k, cap), d_id.numpy(k * cap * 3).reshape(k, cap, 3),
            d_n.numpy(k))


@dataclass
class DeviceCSR:
    """CSR with 64-bit row pointers held on the device."""
    m: int
    n: int
    nnz: int
    row_ptr: DeviceArray
    col: DeviceArray
    val: DeviceArray
    channels: object = None  # device list of the rows' channel ids, if known

    def to_host(self):
        return (self.row_ptr.numpy(self.m + 1), self.col.numpy(self.nnz),
                self.val.numpy(self.nnz))

    @property
    def bytes(self):
        return self.nnz * (self.val.dtype.itemsize + 4) + (self.m + 1) * 8


DEFAULT_STAGING = 64 << 20  # staging entries of the chunked build


def build_csr(setup, grid, cnl_list, nrays, seed=1, raygen=capi.RAYGEN_RAND, stream=None,
              staging=DEFAULT_STAGING, capacity=None, d_list=None, keep_zeros=False):
    """Matrix of the listed channels (rows in list order), zeros dropped
    (keep_zeros: every culled candidate stored, as the reference's fused
    kernel does; raygen RAYGEN_RAND_ELEMENT: rays per element).
    capacity: None = exact count pass first; "auto" = estimate from a count
    pass over every 64th row (falls back to the exact pass if exceeded);
    an int = caller's bound."""
    flags = capi.SM_KEEP_ZEROS if keep_zeros else 0
    lst = np.ascontiguousarray(cnl_list, np.int32)
    m = len(lst)
    if d_list is None:
        d_list = DeviceArray.from_numpy(lst)
    s, g = setup.struct(), grid.struct()
    L = capi.lib()
    ws = Workspace(L.raptr_sm_build_workspace_size(C.byref(g), int(nrays), int(staging)))
    rp = DeviceArray(m + 1, np.int64)
    nnz = C.c_int64()
    if capacity == "auto":
        if m < 4096:
            capacity = None
        else:
            sample = lst[::64].copy()
            d_s = DeviceArray.from_numpy(sample)
            rps = DeviceArray(len(sample) + 1, np.int64)
            capi.call("raptr_sm_build_csr_ex", C.byref(s), C.byref(g), d_s.ptr, len(sample),
                      int(nrays), int(seed), int(raygen), flags, rps.ptr, None, None, 0, C.byref(nnz),
                      ws.ptr, ws.nbytes, stream)
            est = nnz.value * m / len(sample)
            capacity = int(est * 1.03) + 65536
            col = DeviceArray(capacity, np.int32)
            val = DeviceArray(capacity, np.float32)
            st = L.raptr_sm_build_csr_ex(C.byref(s), C.byref(g), d_list.ptr, m, int(nrays),
                                         int(seed), int(raygen), flags, rp.ptr, col.ptr, val.ptr,
                                      int(capacity), C.byref(nnz), ws.ptr, ws.nbytes, stream)
            if st == capi.RAPTR_OK:
                return DeviceCSR(m, grid.nvox, nnz.value, rp, col, val, d_list)
            if st != capi.RAPTR_ERR_WORKSPACE:
                capi.check(st, "raptr_sm_build_csr_ex")
            del col, val
            capacity = None
    if capacity is None:
        capi.call("raptr_sm_build_csr_ex", C.byref(s), C.byref(g), d_list.ptr, m, int(nrays),
                  int(seed), int(raygen), flags, rp.ptr, None, None, 0, C.byref(nnz), ws.ptr,
                  ws.nbytes, stream)
        capacity = nnz.value
    col = DeviceArray(max(capacity, 1), np.int32)
    val = DeviceArray(max(capacity, 1), np.float32)
    capi.call("raptr_sm_build_csr_ex", C.byref(s), C.byref(g), d_list.ptr, m, int(nrays),
              int(seed), int(raygen), flags, rp.ptr, col.ptr, val.ptr, int(capacity), C.byref(nnz),
              ws.ptr, ws.nbytes, stream)
    return DeviceCSR(m, grid.nvox, nnz.value, rp, col, val, d_list)


# midpoint tile of the default locality order
LOCALITY_TILE = 8


ORDERS = {"angle": 0, "y": 1, "slope": 2, "angle_slope": 3}


def locality_order(setup, channels, stream=None, tile=LOCALITY_TILE, kind="angle"):
    """Plan row order that puts lines through the same voxels next to each
    other (raptr_locality_order_ex): positions into ``channels`` sorted by
    angle, midpoint (in tile x tile tiles) and slope (kind "angle"), by
    the midpoint's y first (kind "y"), with the slopes in 4 x 4 tiles (kind
    "slope"), or that over angles in groups of 4, the angle within its group
    last (kind "angle_slope").  Returns a device int32 array."""
    if isinstance(channels, DeviceArray):
        d_ch, m = channels, channels.count
    else:
        lst = np.ascontiguousarray(channels, np.int32)
        d_ch, m = DeviceArray.from_numpy(lst, stream=stream), len(lst)
    out = DeviceArray(max(m, 1), np.int32)
    s = setup.struct()
    capi.call("raptr_locality_order_ex", C.byref(s), d_ch.ptr, int(m), int(tile),
              ORDERS[kind], out.ptr, stream)
    return out


def sensitivity(setup, grid, nrays, seed=1, raygen=capi.RAYGEN_RAND, stream=None):
    """FP64 column sums over all channels of the setup (device array)."""
    out = DeviceArray(grid.nvox, np.float64)
    s, g = setup.struct(), grid.struct()
    capi.call("raptr_sensitivity_f64", C.byref(s), C.byref(g), int(nrays), int(seed),
              int(raygen), out.ptr, stream)
    return out


def backproject(setup, grid, nrays, channels, counts, seed=1, raygen=capi.RAYGEN_RAND,
                normalise=True, stream=None):
    """Summed back-projection A^T y of the measured channels (FP64,
    matrix-free), divided by its sum when normalise (the reference's
    backprojection driver, tests/backprojection.cu:144-185)."""
    ch = np.ascontiguousarray(channels, np.int32)
    w = np.ascontiguousarray(counts, np.float32)
    if len(ch) != len(w):
        raise ValueError("channels and counts differ in length")
    out = DeviceArray(grid.nvox, np.float64)
    d_c = DeviceArray.from_numpy(ch if len(ch) else np.zeros(1, np.int32))
    d_w = DeviceArray.from_numpy(w if len(w) else np.zeros(1, np.float32))
    s, g = setup.struct(), grid.struct()
    capi.call("raptr_backproject_f64", C.byref(s), C.byref(g), d_c.ptr, d_w.ptr, len(ch),
              int(nrays), int(seed), int(raygen), out.ptr, stream)
    x = out.numpy(stream=stream)
    if normalise:
        tot = x.sum()
        x = x / tot if tot != 0 else x
    return x


def forward_project(setup, grid, x, nrays, cnl_list=None, seed=1, raygen=capi.RAYGEN_RAND,
                    stream=None):
    """FP64 projection of image x onto the listed channels (all if None),
    matrix-free.  Returns a host float64 array."""
    dx = x if isinstance(x, DeviceArray) else DeviceArray.from_numpy(x, np.float32)
    if cnl_list is None:
        d_list, m = None, setup.nchannels
    else:
        lst = np.ascontiguousarray(cnl_list, np.int32)
        d_list, m = DeviceArray.from_numpy(lst), len(lst)
    out = DeviceArray(max(m, 1), np.float64)
    s, g = setup.struct(), grid.struct()
    capi.call("raptr_forward_project_f64", C.byref(s), C.byref(g),
              d_list.ptr if d_list is not None else None, m, int(nrays), int(seed),
              int(raygen), dx.ptr, out.ptr, stream)
    return out.numpy(m)
