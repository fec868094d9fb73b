"""ctypes binding of liblt_amd.so (declarations: include/lt_amd.h).

The library is loaded from libtensor_amd/_lib/; a missing library is an
error (there is no fallback implementation).
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "liblt_amd.so")

c_size_p = ctypes.POINTER(ctypes.c_size_t)
c_double_p = ctypes.POINTER(ctypes.c_double)
c_int_p = ctypes.POINTER(ctypes.c_int)
vp = ctypes.c_void_p

# GEMM kernel configurations (lt_btod_contract2_set_tile: 0 .. GEMM_CONFIGS - 1)
GEMM_CONFIGS = 28
# configurations from here on are the stacked kernels (measured by the autotuner)
GEMM_STACKED_FIRST = 12
LT_MODE_PER_BLOCK = 0
LT_MODE_GROUPED = 1
LT_MODE_OZAKI = 2


class LtStats(ctypes.Structure):
    _fields_ = [
        ("t_schedule_ms", ctypes.c_double),
        ("t_plan_ms", ctypes.c_double),
        ("t_gpu_ms", ctypes.c_double),
        ("t_gemm_ms", ctypes.c_double),
        ("flops", ctypes.c_double),
        ("bytes_perm", ctypes.c_double),
        ("n_cblocks", ctypes.c_uint64),
        ("n_pairs", ctypes.c_uint64),
        ("n_tiles", ctypes.c_uint64),
        ("n_unfold", ctypes.c_uint64),
        ("n_launches", ctypes.c_uint64),
        ("tile_cfg", ctypes.c_int32),
        ("ozaki_slices", ctypes.c_int32),
        ("ozaki_max_rel_err", ctypes.c_double),
        ("tune_ms", ctypes.c_double * 32),
        ("graph", ctypes.c_int32),
        ("tile_order", ctypes.c_int32),
        ("n_plans", ctypes.c_uint64),
        ("n_batches", ctypes.c_uint64),
    ]


class LtRequestStats(ctypes.Structure):
    _fields_ = [
        ("t_queue_ms", ctypes.c_double),
        ("t_run_ms", ctypes.c_double),
        ("t_total_ms", ctypes.c_double),
        ("batch", ctypes.c_uint64),
        ("batch_size", ctypes.c_uint64),
        ("status", ctypes.c_int32),
        ("pad", ctypes.c_int32),
    ]


# name -> (argtypes)
_SIGS = {
    "lt_version": [],
    "lt_ctx_create": [ctypes.c_int, ctypes.c_size_t, ctypes.POINTER(vp)],
    "lt_ctx_destroy": [vp],
    "lt_ctx_sync": [vp],
    "lt_ctx_mem": [vp, c_size_p, c_size_p],
    "lt_ctx_set_cap": [vp, ctypes.c_size_t],
    "lt_ctx_set_threads": [vp, ctypes.c_size_t],
    "lt_ctx_last_op_ms": [vp, c_double_p],
    "lt_dev_alloc": [vp, ctypes.c_size_t, ctypes.POINTER(vp)],
    "lt_dev_free": [vp, vp],
    "lt_dev_upload": [vp, vp, vp, ctypes.c_size_t],
    "lt_dev_download": [vp, vp, vp, ctypes.c_size_t],
    "lt_tod_copy": [vp, ctypes.c_size_t, c_size_p, vp, c_size_p, ctypes.c_double,
                    ctypes.c_int, vp],
    "lt_tod_set": [vp, ctypes.c_size_t, ctypes.c_double, vp],
    "lt_tod_add": [vp, ctypes.c_size_t, c_size_p, ctypes.c_size_t, ctypes.POINTER(vp),
                   c_size_p, c_double_p, ctypes.c_int, vp],
    "lt_tod_contract2": [vp, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_size_t,
                         c_size_p, c_size_p, c_size_p, c_size_p, vp, c_size_p, vp,
                         ctypes.c_double, ctypes.c_int, vp],
    "lt_tod_contract2_noalloc": [vp, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_size_t,
                                 c_size_p, c_size_p, c_size_p, c_size_p, vp, c_size_p, vp,
                                 ctypes.c_double, ctypes.c_int, vp],
    "lt_tod_dotprod": [vp, ctypes.c_size_t, c_size_p, vp, c_size_p, c_size_p, vp,
                       c_size_p, ctypes.c_double, c_double_p],
    "lt_product_table_create": [ctypes.c_char_p, ctypes.c_size_t],
    "lt_product_table_add_product": [ctypes.c_char_p, ctypes.c_size_t,
                                     ctypes.c_size_t, ctypes.c_size_t],
    "lt_btensor_create": [vp, ctypes.c_size_t, c_size_p, ctypes.POINTER(vp)],
    "lt_btensor_destroy": [vp],
    "lt_btensor_split": [vp, c_int_p, ctypes.c_size_t],
    "lt_btensor_order": [vp, c_size_p],
    "lt_btensor_bidims": [vp, c_size_p],
    "lt_btensor_block_dims": [vp, ctypes.c_size_t, c_size_p],
    "lt_btensor_block_start": [vp, ctypes.c_size_t, c_size_p],
    "lt_btensor_nonzero": [vp, c_size_p, ctypes.c_size_t, c_size_p],
    "lt_btensor_set_block": [vp, ctypes.c_size_t, c_double_p],
    "lt_btensor_get_block": [vp, ctypes.c_size_t, c_double_p],
    "lt_btensor_zero_block": [vp, ctypes.c_size_t],
    "lt_btensor_zero_all": [vp],
    "lt_btensor_orbit": [vp, ctypes.c_size_t, c_size_p, c_size_p, c_double_p, c_int_p],
    "lt_btensor_random": [vp, ctypes.c_uint64, ctypes.c_int, ctypes.c_double,
                          ctypes.c_int, ctypes.c_int],
    "lt_btensor_export": [vp, c_double_p],
    "lt_btensor_import": [vp, c_double_p],
    "lt_btensor_copy_symmetry": [vp, vp],
    "lt_btensor_set_immutable": [vp],
    "lt_btensor_is_immutable": [vp, c_int_p],
    "lt_btensor_merge_symmetry": [vp, vp],
    "lt_se_perm_create": [ctypes.c_size_t, c_size_p, ctypes.c_double, ctypes.POINTER(vp)],
    "lt_se_part_create": [vp, c_size_p, ctypes.POINTER(vp)],
    "lt_se_part_add_map": [vp, c_size_p, c_size_p, ctypes.c_double],
    "lt_se_part_mark_forbidden": [vp, c_size_p],
    "lt_se_label_create": [vp, ctypes.c_char_p, ctypes.POINTER(vp)],
    "lt_se_label_assign": [vp, c_int_p, ctypes.c_size_t, ctypes.c_size_t],
    "lt_se_label_set_rule": [vp, c_size_p, ctypes.c_size_t],
    "lt_se_label_clear_rule": [vp],
    "lt_se_label_add_product": [vp, ctypes.c_size_t, c_size_p, c_size_p],
    "lt_symel_destroy": [vp],
    "lt_btensor_insert": [vp, vp],
    "lt_btod_contract2_create": [ctypes.c_size_t, ctypes.c_size_t, ctypes.c_size_t,
                                 c_size_p, c_size_p, c_size_p, vp, ctypes.c_double,
                                 vp, ctypes.c_double, ctypes.c_double,
                                 ctypes.POINTER(vp)],
    "lt_btod_contract2_destroy": [vp],
    "lt_btod_contract2_perform": [vp, vp, ctypes.c_int],
    "lt_btod_contract2_perform_add": [vp, vp, ctypes.c_double, ctypes.c_int],
    "lt_btod_contract2_schedule": [vp, c_size_p, ctypes.c_size_t, c_size_p],
    "lt_btod_contract2_get_symmetry": [vp, vp],
    "lt_btod_contract2_ozaki_bounds": [vp, c_size_p, c_double_p, c_double_p, ctypes.c_size_t,
                                       c_size_p],
    "lt_btod_contract2_symc_orbit": [vp, ctypes.c_size_t, c_size_p, c_size_p,
                                     c_double_p, c_int_p],
    "lt_btod_contract2_clst": [vp, ctypes.c_size_t, c_double_p, ctypes.c_size_t, c_size_p],
    "lt_btod_contract2_compute_block_host": [vp, ctypes.c_size_t, c_double_p],
    "lt_btod_contract2_compute_block": [vp, ctypes.c_int, ctypes.c_size_t, c_size_p,
                                        ctypes.c_double, vp],
    "lt_btod_contract2_stats": [vp, ctypes.POINTER(LtStats)],
    "lt_btod_contract2_set_tile": [vp, ctypes.c_int],
    "lt_btod_contract2_set_graph": [vp, ctypes.c_int],
"lt_btod_contract2_set_async": [vp, ctypes.c_int],
    "lt_btod_contract2_set_ozaki": [vp, ctypes.c_int],
    "lt_btod_contract2_set_ozaki_tile": [vp, ctypes.c_int],
    "lt_btod_contract3_create": [ctypes.c_size_t] * 5 + [c_size_p] * 6 + [vp, vp, vp,
                                                                    ctypes.c_double,
                                                                    ctypes.POINTER(vp)],
    "lt_btod_symcontract3_create": [ctypes.c_size_t] * 5 + [c_size_p] * 6 + [
        vp, vp, c_size_p, ctypes.c_int, vp, ctypes.c_double, ctypes.POINTER(vp)],
    "lt_btod_contract3_destroy": [vp],
    "lt_btod_contract3_perform": [vp, vp, ctypes.c_int],
    "lt_btod_contract3_perform_add": [vp, vp, ctypes.c_double, ctypes.c_int],
    "lt_btod_contract3_schedule": [vp, c_size_p, ctypes.c_size_t, c_size_p],
    "lt_btod_contract3_sym_orbit": [vp, ctypes.c_size_t, c_size_p, c_size_p, c_double_p,
                                    c_int_p],
    "lt_btod_contract3_stats": [vp, c_double_p, c_double_p],
    "lt_btod_contract3_plans": [vp, ctypes.POINTER(ctypes.c_uint64)],
    "lt_btod_symmetrize2_create_c3": [vp, c_size_p, ctypes.c_int, ctypes.POINTER(vp)],
    "lt_btod_symmetrize2_create": [vp, c_size_p, ctypes.c_int, ctypes.POINTER(vp)],
    "lt_btod_symmetrize3_create": [vp, c_size_p, c_size_p, ctypes.c_int, ctypes.POINTER(vp)],
    "lt_btod_symmetrize3_create_nested": [vp, c_size_p, c_size_p, ctypes.c_int,
                                          ctypes.POINTER(vp)],
    "lt_btod_symmetrize2_create_nested": [vp, c_size_p, ctypes.c_int, ctypes.POINTER(vp)],
    "lt_btod_symmetrize2_destroy": [vp],
    "lt_btod_symmetrize2_perform": [vp, vp, ctypes.c_int],
    "lt_btod_symmetrize2_perform_add": [vp, vp, ctypes.c_double, ctypes.c_int],
    "lt_btod_symmetrize2_schedule": [vp, c_size_p, ctypes.c_size_t, c_size_p],
    "lt_btod_symmetrize2_sym_orbit": [vp, ctypes.c_size_t, c_size_p, c_size_p,
                                      c_double_p, c_int_p],
    "lt_btod_symmetrize2_stats": [vp, c_double_p, c_double_p, c_double_p, c_double_p],
    "lt_btod_symmetrize2_compute_block": [vp, ctypes.c_int, ctypes.c_size_t, c_size_p,
                                          ctypes.c_double, vp],
    "lt_btod_add": [ctypes.c_size_t, vp, c_size_p, c_double_p, vp, ctypes.c_int,
                    ctypes.c_double],
    "lt_btod_copy": [vp, c_size_p, ctypes.c_double, vp, ctypes.c_int, ctypes.c_double],
    "lt_btod_dotprod": [vp, c_size_p, ctypes.c_double, vp, c_size_p, ctypes.c_double,
                        c_double_p],
    "lt_server_create": [vp, ctypes.c_size_t, ctypes.c_size_t, ctypes.POINTER(vp)],
    "lt_server_destroy": [vp],
    "lt_server_submit": [vp, vp, vp, ctypes.c_int, ctypes.c_double,
                         ctypes.POINTER(ctypes.c_uint64)],
    "lt_server_wait": [vp, ctypes.c_uint64, ctypes.POINTER(LtRequestStats)],
    "lt_server_stats": [vp, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64),
                        c_double_p, ctypes.POINTER(ctypes.c_uint64)],
}

_OPTIONAL = {}

_lib = None


class LtError(RuntimeError):
    """Error returned by the C ABI (code and message)."""

    def __init__(self, code, msg):
        RuntimeError.__init__(self, "lt error %d: %s" % (code, msg))
        self.code = code


def lib():
    """Loads liblt_amd.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("liblt_amd.so not built: run `python -m libtensor_amd.build` "
                              "(expected at %s)" % LIB_PATH)
        l = ctypes.CDLL(LIB_PATH)
        for name, args in _SIGS.items():
            f = getattr(l, name)
            f.argtypes = args
            f.restype = ctypes.c_int
        l.lt_last_error.argtypes = []
        l.lt_last_error.restype = ctypes.c_char_p
        _lib = l
    return _lib


def declared_symbols():
    return sorted(list(_SIGS.keys()) + ["lt_last_error"])


def check(rc):
    if rc != 0:
        raise LtError(rc, lib().lt_last_error().decode("utf-8", "replace"))
    return rc


def call(name, *args):
    return check(getattr(lib(), name)(*args))


def sizes(seq):
    seq = [int(x) for x in seq]
    return (ctypes.c_size_t * max(1, len(seq)))(*seq)


def ints(seq):
    seq = [int(x) for x in seq]
    return (ctypes.c_int * max(1, len(seq)))(*seq)


def dptr(arr):
    """double* of a C-contiguous float64 numpy array."""
    return arr.ctypes.data_as(c_double_p)
