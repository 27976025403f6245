"""ctypes declarations of the C ABI in ``include/raptr_hip.h``.

The shared library is ``raptr_amd/lib/libraptr_hip.so`` (built in-tree by
``raptr_amd/build.py``).  There is no fallback: importing the product API
without the library raises immediately.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libraptr_hip.so")

RAPTR_OK = 0
RAPTR_ERR_INVALID_ARG = 1
RAPTR_ERR_DEVICE = 2
RAPTR_ERR_OUT_OF_MEMORY = 3
RAPTR_ERR_BUDGET = 4
RAPTR_ERR_WORKSPACE = 5
RAPTR_ERR_NOT_SUPPORTED = 6

RAYGEN_RAND = 0
RAYGEN_REGU = 1
RAYGEN_RAND_ELEMENT = 2
SM_KEEP_ZEROS = 1
OP_NON_TRANSPOSE = 0
OP_TRANSPOSE = 1
MAX_ANGLES = 360


class raptr_setup_t(C.Structure):
    _fields_ = [("pos0x", C.c_float), ("pos1x", C.c_float), ("na", C.c_int),
                ("n0z", C.c_int), ("n0y", C.c_int), ("n1z", C.c_int), ("n1y", C.c_int),
                ("da", C.c_float), ("segx", C.c_float), ("segy", C.c_float),
                ("segz", C.c_float)]


class raptr_grid_t(C.Structure):
    _fields_ = [("o", C.c_float * 3), ("d", C.c_float * 3), ("n", C.c_int * 3)]


class raptr_recon_options(C.Structure):
    _fields_ = [("layout", C.c_int), ("d_row_order", C.c_void_p),
                ("grid", C.POINTER(raptr_grid_t)), ("group", C.c_int),
                ("grouped_only", C.c_int)]


class RaptrError(RuntimeError):
    def __init__(self, status, where, message):
        super().__init__(f"{where}: status {status}: {message}")
        self.status = status


P = C.POINTER
vp = C.c_void_p
i32, i64, u32, f32, f64, sz = C.c_int, C.c_int64, C.c_uint32, C.c_float, C.c_double, C.c_size_t
SP, GP = P(raptr_setup_t), P(raptr_grid_t)

# name -> (restype, argtypes); every status-returning function returns int
_SIGS = {
    "raptr_status_string": (C.c_char_p, [i32]),
    "raptr_last_error": (C.c_char_p, []),
    "raptr_version": (i32, []),
    "raptr_device_count": (i32, [P(i32)]),
    "raptr_set_device": (i32, [i32]),
    "raptr_device_name": (i32, [C.c_char_p, i32]),
    "raptr_malloc": (i32, [P(vp), sz]),
    "raptr_free": (i32, [vp]),
    "raptr_memcpy_h2d": (i32, [vp, vp, sz, vp]),
    "raptr_memcpy_d2h": (i32, [vp, vp, sz, vp]),
    "raptr_memcpy_d2d": (i32, [vp, vp, sz, vp]),
    "raptr_memset": (i32, [vp, i32, sz, vp]),
    "raptr_stream_copy": (i32, [vp, vp, i64, vp]),
    "raptr_stream_create": (i32, [P(vp)]),
    "raptr_stream_destroy": (i32, [vp]),
    "raptr_stream_sync": (i32, [vp]),
    "raptr_device_sync": (i32, []),
    "raptr_event_create": (i32, [P(vp)]),
    "raptr_event_destroy": (i32, [vp]),
    "raptr_event_record": (i32, [vp, vp]),
    "raptr_event_elapsed_ms": (i32, [P(f32), vp, vp]),
    "raptr_mem_info": (i32, [P(sz), P(sz)]),
    "raptr_angle_table": (i32, [SP, P(f32), P(f32)]),
    "raptr_sum_radii": (f32, [SP, GP]),
    "raptr_workqueue_workspace_size": (sz, [GP, i32]),
    "raptr_workqueue": (i32, [SP, GP, vp, i32, P(i32), P(i32), i64, vp, vp, i64, P(i64), vp,
                              sz, vp]),
    "raptr_workqueue_counts": (i32, [SP, GP, vp, i32, vp, vp]),
    "raptr_sm_from_wq": (i32, [SP, GP, vp, vp, vp, i64, i32, u32, i32, vp]),
    "raptr_generate_rays": (i32, [SP, vp, i32, i32, u32, i32, vp, vp]),
    "raptr_get_chords": (i32, [vp, i32, GP, i32, vp, vp, vp, vp]),
    "raptr_sm_build_workspace_size": (sz, [GP, i32, i64]),
    "raptr_sm_build_csr": (i32, [SP, GP, vp, i32, i32, u32, i32, vp, vp, vp, i64, P(i64), vp,
                                 sz, vp]),
    "raptr_sm_build_csr_ex": (i32, [SP, GP, vp, i32, i32, u32, i32, i32, vp, vp, vp, i64,
                                    P(i64), vp, sz, vp]),
    "raptr_sm_build_csr_noalloc": (i32, [SP, GP, vp, i32, i32, u32, i32, i32, vp, vp, vp, i64,
                                         P(i64), vp]),
    "raptr_sensitivity_f64": (i32, [SP, GP, i32, u32, i32, vp, vp]),
    "raptr_backproject_f64": (i32, [SP, GP, vp, vp, i32, i32, u32, i32, vp, vp]),
    "raptr_forward_project_f64": (i32, [SP, GP, vp, i32, i32, u32, i32, vp, vp, vp]),
    "raptr_coo_sort_workspace_size": (sz, [i64]),
    "raptr_coo_sort": (i32, [vp, vp, vp, i64, vp, sz, vp]),
    "raptr_coo2csr": (i32, [vp, vp, i64, i32, vp]),
    "raptr_csr2ecsr": (i32, [vp, vp, i32, vp, i32, vp]),
    "raptr_sm_index_arrays": (i32, [vp, vp, i32, i64, i32, vp, vp, vp, vp]),
    "raptr_csr_transpose_workspace_size": (sz, [i64, i32, i32]),
    "raptr_csr_transpose": (i32, [i32, i32, i64, vp, vp, vp, vp, vp, vp, vp, sz, vp]),
    "raptr_csrmv_f32": (i32, [i32, i32, i32, i64, f32, vp, vp, vp, vp, f32, vp, vp]),
    "raptr_csrmv_f64": (i32, [i32, i32, i32, i64, f64, vp, vp, vp, vp, f64, vp, vp]),
    "raptr_spmv_csr64": (i32, [i32, vp, vp, vp, vp, vp, i32, vp]),
    "raptr_divides": (i32, [vp, vp, vp, i32, vp]),
    "raptr_divides_multiplies": (i32, [vp, vp, vp, vp, i32, vp]),
    "raptr_scales": (i32, [vp, f32, i32, vp]),
    "raptr_sum_workspace_size": (sz, [i32]),
    "raptr_sum": (i32, [vp, i32, P(f32), vp, sz, vp]),
    "raptr_column_sums_f64": (i32, [i32, vp, vp, vp, vp]),
    "raptr_mlem_workspace_size": (sz, [i32, i32]),
    "raptr_mlem_forward_ratio": (i32, [i32, vp, vp, vp, vp, vp, vp, i32, vp]),
    "raptr_mlem_backward_update": (i32, [i32, vp, vp, vp, vp, vp, vp, vp, vp, i32, vp]),
    "raptr_mlem": (i32, [i32, i32, vp, vp, vp, vp, vp, vp, vp, vp, vp, i32, i32, i32, i32, vp,
                         sz, vp]),
    "raptr_mlem_autotune": (i32, [i32, i32, vp, vp, vp, vp, vp, vp, vp, vp, i32, P(i32), P(i32),
                                  P(f32), P(f32), vp, sz, vp]),
    "raptr_mlem_partial_slots": (i32, [i32]),
    "raptr_recon_create": (i32, [P(vp), i32, i32, i64, vp, vp, vp, vp, i32, i32, i32, vp]),
    "raptr_recon_destroy": (i32, [vp]),
    "raptr_recon_info": (i32, [vp, P(i64), P(i64), P(i32), P(i32), P(i32), P(i32)]),
    "raptr_recon_set_config": (i32, [vp, i32, i32, i32]),
    "raptr_recon_set_backward": (i32, [vp, i32, i32, i32]),
    "raptr_recon_create_ex": (i32, [vp, i32, i32, i64, vp, vp, vp, vp, i32, i32, i32, i32, vp]),
    "raptr_recon_create_ordered": (i32, [vp, i32, i32, i64, vp, vp, vp, vp, i32, i32, i32, i32,
                                         vp, vp]),
    "raptr_recon_row_ordered": (i32, [vp]),
    "raptr_recon_set_compact_rows": (i32, [vp, i32]),
    "raptr_recon_get_compact_rows": (i32, [vp]),
    "raptr_recon_set_forward_kind": (i32, [vp, i32]),
    "raptr_recon_get_forward_kind": (i32, [vp, P(i32), P(i32), P(i64)]),
    "raptr_recon_set_fused": (i32, [vp, i32, i32, i32, i32]),
    "raptr_recon_clone": (i32, [vp, P(vp), vp]),
    "raptr_recon_clone_count": (i32, [vp]),
    "raptr_stream_wait_event": (i32, [vp, vp]),
    "raptr_recon_get_fused": (i32, [vp, P(i32), P(i32), P(i32), P(i32), P(i32)]),
    "raptr_recon_set_fused_reduce": (i32, [vp, i32]),
    "raptr_recon_set_fused_steps": (i32, [vp, i32]),
    "raptr_recon_get_fused_steps": (i32, [vp]),
    "raptr_recon_get_fused_reduce": (i32, [vp]),
    "raptr_recon_set_fused_tiles": (i32, [vp, i32, i32, vp]),
    "raptr_recon_set_union_rows": (i32, [vp, i32, vp]),
    "raptr_recon_get_unions": (i32, [vp, P(i32), P(i32), P(i64)]),
    "raptr_recon_get_fused_tiles": (i32, [vp, P(i32), P(i32), P(i64), P(i32), P(i32)]),
    "raptr_recon_fused_pass": (i32, [vp, vp, vp, i32, vp]),
    "raptr_locality_order": (i32, [vp, vp, i32, i32, vp, vp]),
    "raptr_locality_order_ex": (i32, [vp, vp, i32, i32, i32, vp, vp]),
    "raptr_recon_layout": (i32, [vp]),
    "raptr_recon_set_blocks_per_launch": (i32, [vp, i32]),
    "raptr_recon_get_blocks_per_launch": (i32, [vp]),
    "raptr_recon_transpose_nnz": (i64, [vp]),
    "raptr_recon_set_broadcast": (i32, [vp, i32]),
    "raptr_recon_graph_count": (i32, [vp]),
    "raptr_recon_set_xcd_blocks": (i32, [vp, i32]),
    "raptr_recon_get_xcd_blocks": (i32, [vp]),
    "raptr_recon_get_broadcast": (i32, [vp]),
    "raptr_recon_set_grouped": (i32, [vp, i32, i32, i32]),
    "raptr_recon_get_grouped": (i32, [vp, vp, vp, vp, vp, vp, vp]),
    "raptr_recon_create_opts": (i32, [vp, i32, i32, i64, vp, vp, vp, vp, i32, i32, i32, vp, vp]),
    "raptr_recon_get_backward": (i32, [vp, vp, vp, vp]),
    "raptr_frame_fill": (i32, [vp, i64, i32, i32, f32, vp]),
    "raptr_frame_scatter": (i32, [vp, i64, i32, i32, vp, vp, i64, vp, vp]),
    "raptr_frame_gather": (i32, [vp, i64, i32, i32, vp, vp]),
    "raptr_frame_copy": (i32, [vp, i64, i32, i32, i32, vp]),
    "raptr_recon_mlem": (i32, [vp, i32, vp, vp, i32, i32, vp]),
    "raptr_recon_mlem_active": (i32, [vp, i32, i32, vp, vp, i32, i32, vp]),
    "raptr_recon_forward": (i32, [vp, i32, vp, vp, vp, vp]),
    "raptr_recon_backward": (i32, [vp, i32, vp, vp]),
}

_lib = None


def lib():
    """The loaded HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"raptr HIP library not found at {LIB_PATH}; build it with "
                "`python -c 'import __graft_entry__ as g; g.build()'` "
                "(there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def exported_symbols():
    return sorted(_SIGS)


def check(status, where):
    if status != RAPTR_OK:
        msg = lib().raptr_last_error().decode(errors="replace")
        raise RaptrError(status, where, msg)


def call(name, *args):
    """Call a status-returning entry point and raise on failure."""
    st = getattr(lib(), name)(*args)
    check(st, name)
    return st
