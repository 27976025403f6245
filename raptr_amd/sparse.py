"""Sparse-format operations on the GPU: cooSort (cooSort.hpp:75-99),
convertCoo2Csr (cusparseWrapper.hpp:53-60), convertCsr2Ecsr
(convertCsr2Ecsr.hpp:51-64), CSRmv (csrmv.hpp:37-99) and the stored
transpose used by the atomic-free back-projection."""
import ctypes as C

import numpy as np

from . import _capi as capi
from .device import DeviceArray, Workspace
from .sysmat import DeviceCSR


def coo_sort(val, row, col, stream=None):
    """Host arrays in, sorted host arrays out (in place on the device)."""
    val = np.ascontiguousarray(val, np.float32)
    row = np.ascontiguousarray(row, np.int32)
    col = np.ascontiguousarray(col, np.int32)
    n = len(val)
    dv, dr, dc = (DeviceArray.from_numpy(a) for a in (val, row, col))
    ws = Workspace(capi.lib().raptr_coo_sort_workspace_size(n))
    capi.call("raptr_coo_sort", dv.ptr, dr.ptr, dc.ptr, n, ws.ptr, ws.nbytes, stream)
    return dv.numpy(n), dr.numpy(n), dc.numpy(n)


def coo2csr(row_id, m, stream=None):
    r = np.ascontiguousarray(row_id, np.int32)
    dr = DeviceArray.from_numpy(r)
    out = DeviceArray(m + 1, np.int32)
    capi.call("raptr_coo2csr", out.ptr, dr.ptr, len(r), int(m), stream)
    return out.numpy()


def csr2ecsr(y_row_id, csr, m, stream=None):
    y = np.ascontiguousarray(y_row_id, np.int32)
    c = np.ascontiguousarray(csr, np.int32)
    dy, dc = DeviceArray.from_numpy(y), DeviceArray.from_numpy(c)
    out = DeviceArray(len(y) + 1, np.int32)
    capi.call("raptr_csr2ecsr", out.ptr, dy.ptr, len(y), dc.ptr, int(m), stream)
    return out.numpy()


def csrmv(op, m, n, alpha, val, row_ptr, col, x, beta, y, dtype=np.float32, stream=None):
    """y = alpha*op(A)*x + beta*y on host arrays (copied through the device)."""
    dt = np.dtype(dtype)
    val = np.ascontiguousarray(val, dt)
    row_ptr = np.ascontiguousarray(row_ptr, np.int32)
    col = np.ascontiguousarray(col, np.int32)
    dv, drp, dcol = (DeviceArray.from_numpy(a) for a in (val, row_ptr, col))
    dx = DeviceArray.from_numpy(np.ascontiguousarray(x, dt))
    dy = DeviceArray.from_numpy(np.ascontiguousarray(y, dt))
    fn = "raptr_csrmv_f32" if dt == np.float32 else "raptr_csrmv_f64"
    capi.call(fn, int(op), int(m), int(n), len(val), alpha, dv.ptr, drp.ptr, dcol.ptr, dx.ptr,
              beta, dy.ptr, stream)
    return dy.numpy()


def transpose(A, stream=None):
    """Stable transpose of a DeviceCSR (rows of A^T keep ascending row order)."""
    L = capi.lib()
    ws = Workspace(L.raptr_csr_transpose_workspace_size(A.nnz, A.m, A.n))
    rp = DeviceArray(A.n + 1, np.int64)
    col = DeviceArray(max(A.nnz, 1), np.int32)
    val = DeviceArray(max(A.nnz, 1), A.val.dtype)
    capi.call("raptr_csr_transpose", A.m, A.n, A.nnz, A.row_ptr.ptr, A.col.ptr, A.val.ptr,
              rp.ptr, col.ptr, val.ptr, ws.ptr, ws.nbytes, stream)
    return DeviceCSR(A.n, A.m, A.nnz, rp, col, val)


def spmv64(A, x, acc64=False, stream=None):
    """y = A x for a DeviceCSR (device x, returns device y)."""
    y = DeviceArray(max(A.m, 1), np.float64 if acc64 else np.float32)
    capi.call("raptr_spmv_csr64", A.m, A.row_ptr.ptr, A.col.ptr, A.val.ptr, x.ptr, y.ptr,
              1 if acc64 else 0, stream)
    return y


def column_sums_f64(At, stream=None):
    out = DeviceArray(At.m, np.float64)
    capi.call("raptr_column_sums_f64", At.m, At.row_ptr.ptr, At.val.ptr, out.ptr, stream)
    return out
