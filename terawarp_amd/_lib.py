"""ctypes declarations of libterawarp.so (include/terawarp.h).

The library is built in-tree by terawarp_amd.build; loading never compiles.
A missing library raises, there is no fallback path.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libterawarp.so")

TW_OK = 0
TW_E_ARG = -1
TW_E_SINGULAR = -2
TW_E_NOMEM = -3
TW_E_DEVICE = -4
TW_E_IO = -5
TW_E_UNSUPPORTED = -6
TW_E_NOCONVERGE = -7

TW_U8 = 1
TW_U16 = 2
TW_MEM_HOST = 0
TW_MEM_DEVICE = 1
TW_PLAIN = 0
TW_MIXED = 1
TW_EXACT = 2
TW_INTERP_LINEAR = 0
TW_INTERP_NN = 1
TW_DF_LINEAR = 0x100
TW_STREAM_BLOCK_ORIGIN = 1
TW_BASIS_BSPLINE = 0
TW_BASIS_LINEAR = 1
TW_DF_F32 = 0
TW_DF_F16 = 1
TW_WARP_TPS = 0
TW_WARP_AFFINE = 1
TW_TILE_AUTO = 0
TW_TILE_GATHER = 1
TW_TILE_LDS = 2
TW_TUNE_TILE = 1
TW_TUNE_TZ = 2
TW_TUNE_FP32_FACTOR_MIN = 3
TW_TUNE_TZ_BATCH = 4
TW_TUNE_SOLVE_WAIT_US = 5
TW_TUNE_SOLVE_PIPELINED = 6

# Every entry point include/terawarp.h declares.
EXPORTS = (
    "tw_create", "tw_destroy", "tw_set_stream", "tw_get_stream", "tw_sync", "tw_set_timing", "tw_set_tile", "tw_set_tune", "tw_get_tune", "tw_get_timings", "tw_last_error",
    "tw_version", "tw_affine_fit", "tw_mode_matrix", "tw_resample_affine", "tw_tps_fit", "tw_tps_grid_eval",
    "tw_grid_release", "tw_resample_tps", "tw_upsample_df", "tw_warp_tps", "tw_warp_tps_noalloc", "tw_warp_batch",
    "tw_resample_tps_bricks", "tw_warp_stream",
    "tw_synth_fill", "tw_synth_read", "tw_checksum64", "tw_checksum_write", "tw_grid_dim",
)
# Every entry point include/terawarp_host.h declares.
HOST_EXPORTS = (
    "tw_v3draw_header", "tw_v3draw_read", "tw_v3draw_write", "tw_markers_read", "tw_markers_write",
    "tw_io_last_error", "tw_plugin_menu_count", "tw_plugin_menu_item", "tw_plugin_func_count",
    "tw_plugin_func_item", "tw_plugin_version", "tw_plugin_parse_params", "tw_plugin_dofunc", "tw_bigimagewarp",
    "tw_tera_block_path", "tw_tera_slice_path", "tw_block_file_write", "tw_stitch_slices", "tw_tera_last_error",
    "tw_slice_dir_open", "tw_slice_dir_size", "tw_slice_dir_read", "tw_slice_dir_close", "tw_downsample_factor",
)

_i64 = ctypes.c_int64
_i32 = ctypes.c_int32
_P = ctypes.c_void_p


class TwVolume(ctypes.Structure):
    _fields_ = [("data", _P), ("nx", _i64), ("ny", _i64), ("nz", _i64), ("nc", _i64),
                ("x0", _i64), ("y0", _i64), ("z0", _i64),
                ("full_nx", _i64), ("full_ny", _i64), ("full_nz", _i64),
                ("dtype", _i32), ("mem", _i32)]


class TwGrid(ctypes.Structure):
    _fields_ = [("data", _P), ("gx", _i64), ("gy", _i64), ("gz", _i64), ("f", _i32), ("storage", _i32),
                ("ox", _i64), ("oy", _i64), ("oz", _i64), ("mem", _i32), ("basis", _i32),
                ("qtab", _P)]


class TwTimings(ctypes.Structure):
    _fields_ = [("fit_ms", ctypes.c_double), ("grid_ms", ctypes.c_double), ("quant_ms", ctypes.c_double),
                ("resample_ms", ctypes.c_double), ("h2d_ms", ctypes.c_double), ("d2h_ms", ctypes.c_double),
                ("total_ms", ctypes.c_double), ("refine_iters", _i32), ("reserved", _i32),
                ("fit_residual", ctypes.c_double)]


class TwWarpItem(ctypes.Structure):
    _fields_ = [("src", TwVolume), ("dst", TwVolume), ("kind", _i32), ("K", _i32),
                ("centres", ctypes.POINTER(ctypes.c_float)), ("targets", ctypes.POINTER(ctypes.c_float)),
                ("M", ctypes.c_double * 12)]


READ_FN = ctypes.CFUNCTYPE(ctypes.c_int, _P, _i64, _i64, _i64, _i64, _i64, _i64, _P)
WRITE_FN = ctypes.CFUNCTYPE(ctypes.c_int, _P, _i64, _i64, _i64, _i64, _i64, _i64, _i64, _P)


class TwStreamDesc(ctypes.Structure):
    _fields_ = [("nx", _i64), ("ny", _i64), ("nz", _i64), ("dtype", _i32), ("brick", _i32), ("f", _i32),
                ("interp", _i32), ("precision", _i32), ("threads", _i32), ("K", _i32), ("flags", _i32),
                ("lam", ctypes.c_double), ("centres", ctypes.POINTER(ctypes.c_float)),
                ("targets", ctypes.POINTER(ctypes.c_float)), ("read", _P), ("read_user", _P),
                ("write", _P), ("write_user", _P)]


class TwStreamStats(ctypes.Structure):
    _fields_ = [("bricks", _i64), ("wall_s", ctypes.c_double), ("setup_ms", ctypes.c_double),
                ("resample_ms", ctypes.c_double), ("read_s", ctypes.c_double), ("write_s", ctypes.c_double),
                ("h2d_bytes", ctypes.c_double), ("d2h_bytes", ctypes.c_double), ("max_margin", ctypes.c_double),
                ("issue_s", ctypes.c_double)]


class TwSynth(ctypes.Structure):
    _fields_ = [("seed", ctypes.c_uint64), ("nx", _i64), ("ny", _i64), ("nz", _i64), ("dtype", _i32),
                ("cell", _i32), ("blobs_per_cell", _i32), ("pad", _i32), ("sigma_lo", ctypes.c_float),
                ("sigma_hi", ctypes.c_float), ("amp_lo", ctypes.c_float), ("amp_hi", ctypes.c_float),
                ("noise", ctypes.c_float)]


class TwChecksumSink(ctypes.Structure):
    _fields_ = [("sums", ctypes.POINTER(ctypes.c_uint64)), ("nbricks", _i64), ("bytes_per_voxel", _i32),
                ("nkeep", _i32), ("keep_ids", ctypes.POINTER(_i64)), ("keep_ptrs", ctypes.POINTER(_P))]


class TwBlockFiles(ctypes.Structure):
    _fields_ = [("base", ctypes.c_char_p), ("dtype", _i32), ("pad", _i32)]


class TwWarpParams(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("device_mode", ctypes.c_int), ("resize", ctypes.c_int),
                ("downsample", ctypes.c_int), ("para", ctypes.c_double * 3), ("newsize", _i64 * 3),
                ("legacy", ctypes.c_int)]


_lib = None


def load():
    """Load libterawarp.so and declare its signatures."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built; run `python -m terawarp_amd.build` (no CPU fallback exists)")
    # torch ships its own libamdhip64.so.7 / libhsa-runtime64.so.1 under the
    # same sonames as /opt/rocm's, and the first one loaded serves the whole
    # process.  Import torch first so that both torch and this library run on
    # torch's runtime (loaded the other way round, torch's device discovery
    # and the library's both fail).
    import torch  # noqa: F401
    L = ctypes.CDLL(LIB_PATH)
    V, G, T = ctypes.POINTER(TwVolume), ctypes.POINTER(TwGrid), ctypes.POINTER(TwTimings)
    dp = ctypes.POINTER(ctypes.c_double)
    fp = ctypes.POINTER(ctypes.c_float)
    sig = {
        "tw_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_P)]),
        "tw_destroy": (ctypes.c_int, [_P]),
        "tw_set_stream": (ctypes.c_int, [_P, _P]),
        "tw_get_stream": (_P, [_P]),
        "tw_sync": (ctypes.c_int, [_P]),
        "tw_set_timing": (ctypes.c_int, [_P, ctypes.c_int]),
        "tw_set_tile": (ctypes.c_int, [_P, ctypes.c_int]),
        "tw_set_tune": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int64]),
        "tw_get_tune": (ctypes.c_int, [_P, ctypes.c_int, ctypes.POINTER(ctypes.c_int64)]),
        "tw_get_timings": (ctypes.c_int, [_P, T]),
        "tw_last_error": (ctypes.c_char_p, [_P]),
        "tw_version": (ctypes.c_char_p, []),
        "tw_affine_fit": (ctypes.c_int, [fp, fp, ctypes.c_int, dp]),
        "tw_mode_matrix": (ctypes.c_int, [ctypes.c_int, dp, _i64, _i64, _i64, dp]),
        "tw_resample_affine": (ctypes.c_int, [_P, V, V, dp, ctypes.c_int]),
        "tw_tps_fit": (ctypes.c_int, [_P, fp, fp, ctypes.c_int, ctypes.c_double, ctypes.c_int, dp, dp]),
        "tw_tps_grid_eval": (ctypes.c_int, [_P, fp, ctypes.c_int, dp, dp, ctypes.c_int, G]),
        "tw_grid_release": (ctypes.c_int, [_P, G]),
        "tw_resample_tps": (ctypes.c_int, [_P, V, V, G, ctypes.c_int, ctypes.c_int]),
        "tw_upsample_df": (ctypes.c_int, [_P, G, _i64, _i64, _i64, _i64, _i64, _i64, ctypes.c_int, fp,
                                          ctypes.c_int]),
        "tw_warp_tps": (ctypes.c_int, [_P, V, V, fp, fp, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                       ctypes.c_int]),
        "tw_resample_tps_bricks": (ctypes.c_int, [_P, V, V, ctypes.c_int, G, ctypes.c_int, ctypes.c_int]),
        "tw_warp_tps_noalloc": (ctypes.c_int, [_P, V, V, fp, fp, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                ctypes.c_int, ctypes.c_int]),
        "tw_warp_batch": (ctypes.c_int, [_P, ctypes.POINTER(TwWarpItem), ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                         ctypes.c_int, ctypes.c_int, ctypes.POINTER(_i32), dp]),
        "tw_warp_stream": (ctypes.c_int, [_P, ctypes.POINTER(TwStreamDesc), ctypes.POINTER(TwStreamStats)]),
        "tw_synth_fill": (ctypes.c_int, [ctypes.POINTER(TwSynth), _i64, _i64, _i64, _i64, _i64, _i64, _P]),
        "tw_synth_read": (ctypes.c_int, [_P, _i64, _i64, _i64, _i64, _i64, _i64, _P]),
        "tw_checksum64": (ctypes.c_uint64, [_P, _i64]),
        "tw_checksum_write": (ctypes.c_int, [_P, _i64, _i64, _i64, _i64, _i64, _i64, _i64, _P]),
        "tw_grid_dim": (_i64, [_i64, ctypes.c_int]),
        "tw_tera_block_path": (ctypes.c_int, [ctypes.c_char_p, _i64, _i64, _i64, ctypes.c_char_p, ctypes.c_int]),
        "tw_tera_slice_path": (ctypes.c_int, [ctypes.c_char_p, _i64, ctypes.c_char_p, ctypes.c_int]),
        "tw_block_file_write": (ctypes.c_int, [_P, _i64, _i64, _i64, _i64, _i64, _i64, _i64, _P]),
        "tw_stitch_slices": (ctypes.c_int, [ctypes.c_char_p, _i64, _i64, _i64, _i32, _i32]),
        "tw_tera_last_error": (ctypes.c_char_p, []),
        "tw_slice_dir_open": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(_P)]),
        "tw_slice_dir_size": (ctypes.c_int, [_P, ctypes.POINTER(_i64), ctypes.POINTER(_i64), ctypes.POINTER(_i64),
                                             ctypes.POINTER(_i32)]),
        "tw_slice_dir_read": (ctypes.c_int, [_P, _i64, _i64, _i64, _i64, _i64, _i64, _P]),
        "tw_slice_dir_close": (None, [_P]),
        "tw_downsample_factor": (ctypes.c_int, [_i64, _i64, _i64]),
        "tw_v3draw_header": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(_i64), ctypes.POINTER(ctypes.c_int),
                                            ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]),
        "tw_v3draw_read": (ctypes.c_int, [ctypes.c_char_p, _P, _i64]),
        "tw_v3draw_write": (ctypes.c_int, [ctypes.c_char_p, _P, ctypes.POINTER(_i64), ctypes.c_int]),
        "tw_markers_read": (ctypes.c_int, [ctypes.c_char_p, fp, ctypes.c_int]),
        "tw_markers_write": (ctypes.c_int, [ctypes.c_char_p, fp, ctypes.c_int]),
        "tw_io_last_error": (ctypes.c_char_p, []),
        "tw_plugin_menu_count": (ctypes.c_int, []),
        "tw_plugin_menu_item": (ctypes.c_char_p, [ctypes.c_int]),
        "tw_plugin_func_count": (ctypes.c_int, []),
        "tw_plugin_func_item": (ctypes.c_char_p, [ctypes.c_int]),
        "tw_plugin_version": (ctypes.c_float, []),
        "tw_plugin_parse_params": (ctypes.c_int, [ctypes.POINTER(ctypes.c_char_p), ctypes.c_int,
                                                  ctypes.POINTER(TwWarpParams)]),
        "tw_plugin_dofunc": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(ctypes.c_char_p), ctypes.c_int,
                                            ctypes.POINTER(ctypes.c_char_p), ctypes.c_int,
                                            ctypes.POINTER(ctypes.c_char_p), ctypes.c_int]),
        "tw_bigimagewarp": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p,
                                           ctypes.POINTER(TwWarpParams), ctypes.c_char_p, ctypes.c_int]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L
