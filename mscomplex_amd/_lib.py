"""ctypes binding of libmscx.so (the C ABI declared in include/mscx.h).

The library must have been built in-tree (python -m mscomplex_amd.build);
there is no fallback implementation, so a missing library raises at import
of the first call.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# MSCX_LIB_PATH selects another build of the same library (the CPU
# sanitizer builds of build.py --sanitize)
LIB_PATH = os.environ.get("MSCX_LIB_PATH") or os.path.join(_HERE, "libmscx.so")

_lib = None

c_int, c_int32, c_uint32, c_uint64, c_size_t, c_double, c_int64 = (
    ctypes.c_int, ctypes.c_int32, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_size_t, ctypes.c_double,
    ctypes.c_int64)
P = ctypes.c_void_p

# name -> (restype, argtypes)
_SIGS = {
    "mscx_init": (c_int, [c_int]),
    "mscx_shutdown": (c_int, []),
    "mscx_last_error": (ctypes.c_char_p, []),
    "mscx_version": (c_int, []),
    "mscx_profile": (c_int, [c_int]),
    "mscx_profile_read": (c_int, [P, P, c_int]),
    "mscx_set_graphs": (c_int, [c_int, P]),
    "mscx_graph_stats": (c_int, [P, P]),
    "mscx_set_block_rows": (c_int, [c_int32, P, P]),
    "mscx_disc_trace_stats": (c_int, [P]),
    "mscx_merge_stats": (c_int, [P]),
    "mscx_dev_alloc": (c_int, [c_size_t, ctypes.POINTER(P)]),
    "mscx_dev_free": (c_int, [P]),
    "mscx_dev_upload": (c_int, [P, P, c_size_t]),
    "mscx_dev_download": (c_int, [P, P, c_size_t]),
    "mscx_tile_create": (c_int, [P, P, ctypes.POINTER(P)]),
    "mscx_tile_destroy": (c_int, [P]),
    "mscx_tile_upload": (c_int, [P, P, c_size_t]),
    "mscx_tile_bind_device": (c_int, [P, P, c_int64]),
    "mscx_tile_work": (c_int, [P]),
    "mscx_tiles_work": (c_int, [P, c_int]),
    "mscx_tile_work_noalloc": (c_int, [P, P, P, c_int64, c_uint32, P, P, P, P, ctypes.POINTER(c_uint32)]),
    "mscx_tile_num_critical": (c_int, [P, ctypes.POINTER(c_uint32)]),
    "mscx_tile_download": (c_int, [P, P, P, P, P]),
    "mscx_tile_download_codes": (c_int, [P, P]),
    "mscx_tile_download_critical": (c_int, [P, P, P, P]),
    "mscx_complex_from_tile": (c_int, [P, ctypes.POINTER(P)]),
    "mscx_complex_from_critical": (c_int, [P, P, c_uint32, P, P, P, P, ctypes.POINTER(P)]),
    "mscx_complex_fill_discs_host": (c_int, [P, P, P, P, P]),
    "mscx_complex_merge_up": (c_int, [P, P, ctypes.POINTER(P)]),
    "mscx_complex_merge_down": (c_int, [P, P, P]),
    "mscx_complex_simplify_un_simplify": (c_int, [P, c_double]),
    "mscx_tile_fill_discs": (c_int, [P, P]),
    "mscx_complex_write_discs": (c_int, [P, ctypes.c_char_p]),
    "mscx_complex_save": (c_int, [P, ctypes.c_char_p]),
    "mscx_complex_load": (c_int, [ctypes.c_char_p, ctypes.POINTER(P)]),
    "mscx_complex_dump": (c_int, [P, P, c_size_t, ctypes.POINTER(c_size_t)]),
    "mscx_complex_num_cps": (c_int, [P, ctypes.POINTER(c_uint32)]),
    "mscx_complex_destroy": (c_int, [P]),
    "mscx_grid_run": (c_int, [P, c_int32, c_int32, c_int32, c_int32, c_int32, c_double, c_uint32,
                              ctypes.POINTER(P)]),
    "mscx_grid_run_file": (c_int, [ctypes.c_char_p, c_int32, c_int32, c_int32, c_int32, c_int32, c_double, c_int32,
                                   c_uint32, ctypes.c_char_p, ctypes.POINTER(P)]),
    "mscx_grid_num_pieces": (c_int, [P, ctypes.POINTER(c_int32)]),
    "mscx_grid_piece": (c_int, [P, c_int32, ctypes.POINTER(P)]),
    "mscx_grid_piece_num_cps": (c_int, [P, c_int32, P]),
    "mscx_grid_checksum": (c_int, [P, P]),
    "mscx_grid_times": (c_int, [P, P]),
    "mscx_grid_destroy": (c_int, [P]),
    "mscx_batch_run": (c_int, [c_int32, P, P, P, P, P, P, c_uint32, ctypes.POINTER(P)]),
    "mscx_batch_size": (c_int, [P, ctypes.POINTER(c_int32)]),
    "mscx_batch_grid": (c_int, [P, c_int32, ctypes.POINTER(P), ctypes.POINTER(c_int32)]),
    "mscx_batch_error": (ctypes.c_char_p, [P, c_int32]),
    "mscx_batch_destroy": (c_int, [P]),
    "mscx_server_create": (c_int, [c_size_t, c_int32, ctypes.POINTER(P)]),
    "mscx_server_set_work_limit": (c_int, [P, c_double]),
    "mscx_server_submit": (c_int, [P, P, c_int32, c_int32, c_int32, c_int32, c_double, c_uint64, c_uint32,
                                   ctypes.POINTER(c_uint64)]),
    "mscx_server_wait": (c_int, [P, c_uint64, c_double, ctypes.POINTER(P), ctypes.POINTER(c_int32),
                                 ctypes.POINTER(c_double)]),
    "mscx_server_error": (ctypes.c_char_p, [P, c_uint64]),
    "mscx_server_release": (c_int, [P, c_uint64]),
    "mscx_server_stats": (c_int, [P, P]),
    "mscx_server_destroy": (c_int, [P]),
    "mscx_gen_gaussians": (c_int, [P, c_int32, c_int32, c_int32, c_double, c_uint64]),
    "mscx_gen_value_noise": (c_int, [P, c_int32, c_int32, c_int32, c_double, c_double, c_double, c_uint64]),
    "mscx_gen_white": (c_int, [P, c_int32, c_int32, c_int32, c_uint64]),
}


class MscxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"mscx error {code}: {msg}")
        self.code = code


def lib():
    """Load libmscx.so once; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run `python -m mscomplex_amd.build` first")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name, None)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def exported_symbols():
    """Names of _SIGS present in the library (for ABI checks)."""
    L = lib()
    return [n for n in _SIGS if hasattr(L, n)]


def check(rc):
    if rc != 0:
        msg = lib().mscx_last_error()
        raise MscxError(rc, msg.decode() if msg else "")
    return rc


def ptr(a):
    """Raw data pointer of a numpy array (None for None)."""
    return None if a is None else ctypes.c_void_p(a.ctypes.data)
