"""The allocation-free matrix build (raptr_sm_build_csr_noalloc, library
scratch instead of a workspace) against raptr_sm_build_csr_ex through the C
ABI: two calls in a row on the same device array, the second with a smaller
input, bit for bit."""
import ctypes as C

import numpy as np
import pytest

from raptr_amd import geometry as G

pytestmark = pytest.mark.gpu


def _noalloc(setup, grid, lst, nrays, raygen, flags):
    from raptr_amd import _capi as capi
    from raptr_amd.device import DeviceArray
    s, g = setup.struct(), grid.struct()
    d_list = DeviceArray.from_numpy(lst)
    rp = DeviceArray(len(lst) + 1, np.int64)
    nnz = C.c_int64()
    capi.call("raptr_sm_build_csr_noalloc", C.byref(s), C.byref(g), d_list.ptr, len(lst),
              int(nrays), 1, int(raygen), int(flags), rp.ptr, None, None, 0, C.byref(nnz), None)
    count = nnz.value
    col = DeviceArray(max(count, 1), np.int32)
    val = DeviceArray(max(count, 1), np.float32)
    capi.call("raptr_sm_build_csr_noalloc", C.byref(s), C.byref(g), d_list.ptr, len(lst),
              int(nrays), 1, int(raygen), int(flags), rp.ptr, col.ptr, val.ptr, count,
              C.byref(nnz), None)
    assert nnz.value == count
    return rp.numpy(len(lst) + 1), col.numpy(count), val.numpy(count)


@pytest.mark.parametrize("raygen,keep", [(0, False), (2, True)])
def test_noalloc_build_equals_workspace_build(gpu, raygen, keep):
    from raptr_amd import sysmat
    setup, grid, nrays = G.config_small()
    rng = np.random.default_rng(3)
    for n_rows in (3000, 400):  # second call: smaller input, same scratch
        lst = np.sort(rng.choice(setup.nchannels, n_rows, replace=False)).astype(np.int32)
        A = sysmat.build_csr(setup, grid, lst, nrays, raygen=raygen, keep_zeros=keep)
        want = A.to_host()
        got = _noalloc(setup, grid, lst, nrays, raygen, 1 if keep else 0)
        for a, b in zip(got, want):
            np.testing.assert_array_equal(a, b)
