"""ctypes bindings of the CPU oracle (test infrastructure only).

Only the test-suite, the smoke check and the benchmark's CPU-baseline leg
import this module.  It loads ``oracle/_build/libraptr_oracle.so`` (the C
restatement) and, when it has been built in a container that holds the raptr
sources, ``oracle/_ref/libraptr_ref.so`` (the reference's own getWorkqueue).
"""
import contextlib
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libraptr_oracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libraptr_ref.so")

RAYGEN_RAND = 0
RAYGEN_REGU = 1
RAYGEN_RAND_ELEMENT = 2
SM_KEEP_ZEROS = 1


class Setup(C.Structure):
    _fields_ = [("pos0x", C.c_float), ("pos1x", C.c_float), ("na", C.c_int),
                ("n0z", C.c_int), ("n0y", C.c_int), ("n1z", C.c_int),
                ("n1y", C.c_int), ("da", C.c_float), ("segx", C.c_float),
                ("segy", C.c_float), ("segz", C.c_float)]


class Grid(C.Structure):
    _fields_ = [("o", C.c_float * 3), ("d", C.c_float * 3), ("n", C.c_int * 3)]


def make_setup(pos0x, pos1x, na, n0z, n0y, n1z, n1y, da, segx, segy, segz):
    return Setup(pos0x, pos1x, na, n0z, n0y, n1z, n1y, da, segx, segy, segz)


def make_grid(o, d, n):
    g = Grid()
    for k in range(3):
        g.o[k], g.d[k], g.n[k] = o[k], d[k], n[k]
    return g


def _p(a, ct):
    return a.ctypes.data_as(C.POINTER(ct)) if a is not None else None


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run make -C oracle)")
        _lib = _bind(C.CDLL(LIB_PATH))
    return _lib


@contextlib.contextmanager
def swap_library(path):
    """Run the bindings against another build of the oracle (the pin tests
    load builds with a planted mistake)."""
    global _lib
    saved = _lib
    _lib = _bind(C.CDLL(path))
    try:
        yield _lib
    finally:
        _lib = saved


def _bind(L):
    P = C.POINTER
    S, G = P(Setup), P(Grid)
    i64, f32, f64, i32, u32 = C.c_int64, C.c_float, C.c_double, C.c_int, C.c_uint32
    sig = {
        "ro_nchannels": (i64, [S]),
        "ro_cnl_split": (None, [S, i32] + [P(i32)] * 5),
        "ro_cnl_lin": (i32, [S, i32, i32, i32, i32, i32]),
        "ro_vxl_split": (None, [G, i32, P(i32), P(i32), P(i32)]),
        "ro_vxl_lin": (i32, [G, i32, i32, i32]),
        "ro_angle_table": (None, [S, P(f32), P(f32)]),
        "ro_trafo": (None, [S, i32, f32, f32, i32, i32, P(f32), P(f32)]),
        "ro_distance": (f32, [P(f32), P(f32), P(f32)]),
        "ro_workqueue": (i64, [S, G, P(i32), i32, P(i32), P(i32), i64, P(i32), P(i32)]),
        "ro_workqueue_counts": (None, [S, G, P(i32), i32, P(i64)]),
        "ro_philox4x32_10": (None, [P(u32), P(u32), P(u32)]),
        "ro_u01": (f32, [u32]),
        "ro_set_threads": (None, [i32]),
        "ro_get_threads": (i32, []),
        "ro_ray": (None, [S, P(f32), P(f32), i32, u32, i32, i32, P(f32)]),
        "ro_intersection_length": (f32, [P(f32), P(f32)]),
        "ro_intersection_length_f64": (f64, [P(f64), P(f64)]),
        "ro_voxel_box": (None, [G, i32, P(f32)]),
        "ro_sm_from_wq": (None, [S, G, P(i32), P(i32), P(f32), i64, i32, u32, i32]),
        "ro_sm_from_wq_f64": (None, [S, G, P(i32), P(i32), P(f64), i64, i32, u32, i32]),
        "ro_get_chords": (i32, [P(f32), P(i32), i32, P(f32), P(f32), P(f32), P(i32)]),
        "ro_build_csr": (i64, [S, G, P(i32), i32, i32, u32, i32, P(i64), P(i32), P(f32), P(f64)]),
        "ro_build_csr_ex": (i64, [S, G, P(i32), i32, i32, u32, i32, i32, P(i64), P(i32), P(f32),
                                  P(f64)]),
        "ro_ray_element": (None, [S, P(f32), P(f32), i32, u32, i32, i32, i32, P(f32)]),
        "ro_coo_sort": (None, [P(f32), P(i32), P(i32), i64]),
        "ro_coo2csr": (None, [P(i32), P(i32), i64, i32]),
        "ro_csr2ecsr": (None, [P(i32), P(i32), i32, P(i32), i32]),
        "ro_csrmv_f32": (None, [i32, i32, i32, f32, P(f32), P(i32), P(i32), P(f32), f32, P(f32)]),
        "ro_csrmv_f64": (None, [i32, i32, i32, f64, P(f64), P(i64), P(i32), P(f64), f64, P(f64)]),
        "ro_divides": (None, [P(f32), P(f32), P(f32), i32]),
        "ro_divides_multiplies": (None, [P(f32), P(f32), P(f32), P(f32), i32]),
        "ro_scales": (None, [P(f32), f32, i32]),
        "ro_sum": (f32, [P(f32), i32]),
        "ro_mlem_f64": (None, [i32, i32, P(i64), P(i32), P(f64), P(f64), P(f64), P(f64), i32]),
        "ro_mlem_f32": (None, [i32, i32, P(i64), P(i32), P(f32), P(f32), P(f64), P(f32), i32]),
        "ro_sensitivity_f64": (None, [i32, i32, P(i64), P(i32), P(f64), P(f64)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    return L


def ref_available():
    return os.path.exists(REF_PATH)


H5REF_PATH = os.path.join(os.path.dirname(REF_PATH), "libraptr_h5ref.so")
_h5ref = None


def h5ref_available():
    return os.path.exists(H5REF_PATH)


def h5ref():
    """The reference's HDF5 writer and readers (oracle/ref/h5_ref.cpp)."""
    global _h5ref
    if _h5ref is None:
        L = C.CDLL(H5REF_PATH)
        P = C.POINTER
        L.ref_h5_write_density.restype = C.c_int
        L.ref_h5_write_density.argtypes = [C.c_char_p, P(C.c_float), P(C.c_float),
                                           P(C.c_float), P(C.c_int)]
        L.ref_h5_read.restype = C.c_int64
        L.ref_h5_read.argtypes = [C.c_char_p, C.c_int, P(C.c_float), C.c_int64]
        _h5ref = L
    return _h5ref


def ref_h5_read(path, density, n):
    out = np.zeros(n, np.float32)
    got = h5ref().ref_h5_read(path.encode(), 1 if density else 0, _p(out, C.c_float), n)
    if got != n:
        raise RuntimeError(f"reference reader failed on {path}")
    return out


def ref():
    global _ref
    if _ref is None:
        L = C.CDLL(REF_PATH)
        P = C.POINTER
        L.ref_workqueue.restype = C.c_int64
        L.ref_workqueue.argtypes = [P(C.c_float), P(C.c_int), P(C.c_float), P(C.c_int),
                                    P(C.c_int), C.c_int, P(C.c_int), P(C.c_int),
                                    C.c_int64, P(C.c_int), P(C.c_int)]
        L.ref_trafo.restype = None
        L.ref_trafo.argtypes = [P(C.c_float), P(C.c_int), C.c_int, C.c_int, C.c_int,
                                C.c_int, P(C.c_float), P(C.c_float)]
        L.ref_distance.restype = C.c_float
        L.ref_distance.argtypes = [P(C.c_float)] * 3
        _ref = L
    return _ref


def _setup_arrays(s):
    f = np.array([s.pos0x, s.pos1x, s.da, s.segx, s.segy, s.segz], np.float32)
    i = np.array([s.na, s.n0z, s.n0y, s.n1z, s.n1y], np.int32)
    return f, i


def _grid_arrays(g):
    f = np.array([g.o[0], g.o[1], g.o[2], g.d[0], g.d[1], g.d[2]], np.float32)
    i = np.array([g.n[0], g.n[1], g.n[2]], np.int32)
    return f, i


# ---------------------------------------------------------------- wrappers
def nvox(g):
    return g.n[0] * g.n[1] * g.n[2]


def nchannels(s):
    return int(lib().ro_nchannels(C.byref(s)))


def angle_table(s):
    sn = np.zeros(s.na, np.float32)
    cs = np.zeros(s.na, np.float32)
    lib().ro_angle_table(C.byref(s), _p(sn, C.c_float), _p(cs, C.c_float))
    return sn, cs


def workqueue_counts(s, g, lst):
    lst = np.ascontiguousarray(lst, np.int32)
    out = np.zeros(len(lst), np.int64)
    lib().ro_workqueue_counts(C.byref(s), C.byref(g), _p(lst, C.c_int), len(lst), _p(out, C.c_int64))
    return out


def workqueue(s, g, lst, list_id=0, vxl_id=0, n=-1, capacity=None):
    """Restated getWorkqueue; returns (cnl, vxl, list_id, vxl_id)."""
    lst = np.ascontiguousarray(lst, np.int32)
    if capacity is None:
        capacity = n if n >= 0 else int(workqueue_counts(s, g, lst[list_id:]).sum())
    cnl = np.zeros(max(capacity, 1), np.int32)
    vxl = np.zeros(max(capacity, 1), np.int32)
    li, vi = C.c_int(list_id), C.c_int(vxl_id)
    k = lib().ro_workqueue(C.byref(s), C.byref(g), _p(lst, C.c_int), len(lst), C.byref(li),
                           C.byref(vi), n, _p(cnl, C.c_int), _p(vxl, C.c_int))
    return cnl[:k].copy(), vxl[:k].copy(), li.value, vi.value


def ref_workqueue(s, g, lst, list_id=0, vxl_id=0, n=-1, capacity=None):
    """The reference's own getWorkqueue (compiled from its sources)."""
    lst = np.ascontiguousarray(lst, np.int32)
    sf, si = _setup_arrays(s)
    gf, gi = _grid_arrays(g)
    if capacity is None:
        capacity = n if n >= 0 else int(workqueue_counts(s, g, lst[list_id:]).sum()) + 1
    cnl = np.zeros(max(capacity, 1), np.int32)
    vxl = np.zeros(max(capacity, 1), np.int32)
    li, vi = C.c_int(list_id), C.c_int(vxl_id)
    k = ref().ref_workqueue(_p(sf, C.c_float), _p(si, C.c_int), _p(gf, C.c_float), _p(gi, C.c_int),
                            _p(lst, C.c_int), len(lst), C.byref(li), C.byref(vi), n,
                            _p(cnl, C.c_int), _p(vxl, C.c_int))
    return cnl[:k].copy(), vxl[:k].copy(), li.value, vi.value


def philox(ctr, key):
    c = np.ascontiguousarray(ctr, np.uint32)
    k = np.ascontiguousarray(key, np.uint32)
    o = np.zeros(4, np.uint32)
    lib().ro_philox4x32_10(_p(c, C.c_uint32), _p(k, C.c_uint32), _p(o, C.c_uint32))
    return o


def ray_element(s, cnl, element, r, seed=1, raygen=RAYGEN_RAND_ELEMENT):
    sn, cs = angle_table(s)
    out = np.zeros(6, np.float32)
    lib().ro_ray_element(C.byref(s), _p(sn, C.c_float), _p(cs, C.c_float), raygen, seed, cnl,
                         element, r, _p(out, C.c_float))
    return out


def ray(s, cnl, r, seed=1, raygen=RAYGEN_RAND):
    sn, cs = angle_table(s)
    o = np.zeros(6, np.float32)
    lib().ro_ray(C.byref(s), _p(sn, C.c_float), _p(cs, C.c_float), raygen, seed, cnl, r,
                 _p(o, C.c_float))
    return o


def trafo(s, det, ida, idz, idy, box):
    sn, cs = angle_table(s)
    b = np.ascontiguousarray(box, np.float32)
    o = np.zeros(3, np.float32)
    lib().ro_trafo(C.byref(s), det, float(sn[ida]), float(cs[ida]), idz, idy, _p(b, C.c_float),
                   _p(o, C.c_float))
    return o


def intersection_length(box, r):
    b = np.ascontiguousarray(box, np.float32)
    rr = np.ascontiguousarray(r, np.float32)
    return float(lib().ro_intersection_length(_p(b, C.c_float), _p(rr, C.c_float)))


def intersection_length_f64(box, r):
    b = np.ascontiguousarray(box, np.float64)
    rr = np.ascontiguousarray(r, np.float64)
    return float(lib().ro_intersection_length_f64(_p(b, C.c_double), _p(rr, C.c_double)))


def voxel_box(g, v):
    o = np.zeros(6, np.float32)
    lib().ro_voxel_box(C.byref(g), v, _p(o, C.c_float))
    return o


def sm_from_wq(s, g, cnl, vxl, nrays, seed=1, raygen=RAYGEN_RAND, f64=False):
    cnl = np.ascontiguousarray(cnl, np.int32)
    vxl = np.ascontiguousarray(vxl, np.int32)
    if f64:
        val = np.zeros(len(cnl), np.float64)
        lib().ro_sm_from_wq_f64(C.byref(s), C.byref(g), _p(cnl, C.c_int), _p(vxl, C.c_int),
                                _p(val, C.c_double), len(cnl), nrays, seed, raygen)
    else:
        val = np.zeros(len(cnl), np.float32)
        lib().ro_sm_from_wq(C.byref(s), C.byref(g), _p(cnl, C.c_int), _p(vxl, C.c_int),
                            _p(val, C.c_float), len(cnl), nrays, seed, raygen)
    return val


def get_chords(r, g, cap=None):
    rr = np.ascontiguousarray(r, np.float32)
    go = np.array(list(g.o), np.float32)
    gd = np.array(list(g.d), np.float32)
    gn = np.array(list(g.n), np.int32)
    cap = cap or int(gn.sum()) + 4
    ch = np.zeros(cap, np.float32)
    ids = np.zeros(3 * cap, np.int32)
    k = lib().ro_get_chords(_p(ch, C.c_float), _p(ids, C.c_int), cap, _p(rr, C.c_float),
                            _p(go, C.c_float), _p(gd, C.c_float), _p(gn, C.c_int))
    return ch[:k].copy(), ids[:3 * k].reshape(-1, 3).copy()


def build_csr(s, g, lst, nrays, seed=1, raygen=RAYGEN_RAND, with_f64=False, flags=0):
    """Matrix of the listed rows, zeros dropped (flags SM_KEEP_ZEROS: kept):
    (row_ptr i64, col, val f32, val f64|None)."""
    lst = np.ascontiguousarray(lst, np.int32)
    rp = np.zeros(len(lst) + 1, np.int64)
    L = lib()
    L.ro_build_csr_ex(C.byref(s), C.byref(g), _p(lst, C.c_int), len(lst), nrays, seed, raygen,
                      flags, _p(rp, C.c_int64), None, None, None)
    nnz = int(rp[-1])
    col = np.zeros(max(nnz, 1), np.int32)
    val = np.zeros(max(nnz, 1), np.float32)
    v64 = np.zeros(max(nnz, 1), np.float64) if with_f64 else None
    L.ro_build_csr_ex(C.byref(s), C.byref(g), _p(lst, C.c_int), len(lst), nrays, seed, raygen,
                      flags, _p(rp, C.c_int64), _p(col, C.c_int), _p(val, C.c_float),
                      _p(v64, C.c_double) if with_f64 else None)
    return rp, col[:nnz], val[:nnz], (v64[:nnz] if with_f64 else None)


def coo_sort(val, row, col):
    val = np.ascontiguousarray(val, np.float32).copy()
    row = np.ascontiguousarray(row, np.int32).copy()
    col = np.ascontiguousarray(col, np.int32).copy()
    lib().ro_coo_sort(_p(val, C.c_float), _p(row, C.c_int), _p(col, C.c_int), len(val))
    return val, row, col


def coo2csr(row_id, m):
    r = np.ascontiguousarray(row_id, np.int32)
    out = np.zeros(m + 1, np.int32)
    lib().ro_coo2csr(_p(out, C.c_int), _p(r, C.c_int), len(r), m)
    return out


def csr2ecsr(y_row_id, csr, m):
    y = np.ascontiguousarray(y_row_id, np.int32)
    c = np.ascontiguousarray(csr, np.int32)
    out = np.zeros(len(y) + 1, np.int32)
    lib().ro_csr2ecsr(_p(out, C.c_int), _p(y, C.c_int), len(y), _p(c, C.c_int), m)
    return out


def csrmv_f32(trans, m, n, alpha, val, rp, col, x, beta, y):
    val = np.ascontiguousarray(val, np.float32)
    rp = np.ascontiguousarray(rp, np.int32)
    col = np.ascontiguousarray(col, np.int32)
    x = np.ascontiguousarray(x, np.float32)
    y = np.ascontiguousarray(y, np.float32).copy()
    lib().ro_csrmv_f32(trans, m, n, alpha, _p(val, C.c_float), _p(rp, C.c_int), _p(col, C.c_int),
                       _p(x, C.c_float), beta, _p(y, C.c_float))
    return y


def csrmv_f64(trans, m, n, alpha, val, rp, col, x, beta, y):
    val = np.ascontiguousarray(val, np.float64)
    rp = np.ascontiguousarray(rp, np.int64)
    col = np.ascontiguousarray(col, np.int32)
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64).copy()
    lib().ro_csrmv_f64(trans, m, n, alpha, _p(val, C.c_double), _p(rp, C.c_int64),
                       _p(col, C.c_int), _p(x, C.c_double), beta, _p(y, C.c_double))
    return y


def divides(a, b):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    r = np.zeros_like(a)
    lib().ro_divides(_p(r, C.c_float), _p(a, C.c_float), _p(b, C.c_float), len(a))
    return r


def divides_multiplies(a, b, c):
    a, b, c = (np.ascontiguousarray(v, np.float32) for v in (a, b, c))
    r = np.zeros_like(a)
    lib().ro_divides_multiplies(_p(r, C.c_float), _p(a, C.c_float), _p(b, C.c_float),
                                _p(c, C.c_float), len(a))
    return r


def scales(x, alpha):
    x = np.ascontiguousarray(x, np.float32).copy()
    lib().ro_scales(_p(x, C.c_float), alpha, len(x))
    return x


def fsum(x):
    x = np.ascontiguousarray(x, np.float32)
    return float(lib().ro_sum(_p(x, C.c_float), len(x)))


def sensitivity_f64(rp, col, val64, nvoxels):
    rp = np.ascontiguousarray(rp, np.int64)
    col = np.ascontiguousarray(col, np.int32)
    v = np.ascontiguousarray(val64, np.float64)
    s = np.zeros(nvoxels, np.float64)
    lib().ro_sensitivity_f64(len(rp) - 1, nvoxels, _p(rp, C.c_int64), _p(col, C.c_int),
                             _p(v, C.c_double), _p(s, C.c_double))
    return s


def mlem_f64(rp, col, val64, y, sens, x0, niter):
    rp = np.ascontiguousarray(rp, np.int64)
    col = np.ascontiguousarray(col, np.int32)
    v = np.ascontiguousarray(val64, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    s = np.ascontiguousarray(sens, np.float64)
    x = np.ascontiguousarray(x0, np.float64).copy()
    lib().ro_mlem_f64(len(rp) - 1, len(x), _p(rp, C.c_int64), _p(col, C.c_int), _p(v, C.c_double),
                      _p(y, C.c_double), _p(s, C.c_double), _p(x, C.c_double), niter)
    return x


def mlem_f32(rp, col, val, y, sens, x0, niter):
    """The recurrence in FP32 (the reference's precision)."""
    rp = np.ascontiguousarray(rp, np.int64)
    col = np.ascontiguousarray(col, np.int32)
    val = np.ascontiguousarray(val, np.float32)
    y = np.ascontiguousarray(y, np.float32)
    sens = np.ascontiguousarray(sens, np.float64)
    x = np.ascontiguousarray(x0, np.float32).copy()
    lib().ro_mlem_f32(len(rp) - 1, len(x), _p(rp, C.c_int64), _p(col, C.c_int),
                      _p(val, C.c_float), _p(y, C.c_float), _p(sens, C.c_double),
                      _p(x, C.c_float), niter)
    return x
