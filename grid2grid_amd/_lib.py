"""ctypes binding of libgrid2grid_amd.so (include/grid2grid_amd.h).

The library must have been built (python -m grid2grid_amd.build); there is
no fallback implementation: importing the bindings without it raises.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgrid2grid_amd.so")

G2G_OK = 0
STATUS = {
    1: "G2G_ERR_INVALID",
    2: "G2G_ERR_HIP",
    3: "G2G_ERR_NOMEM",
    4: "G2G_ERR_BUDGET",
    5: "G2G_ERR_UNSUPPORTED",
    6: "G2G_ERR_STATE",
}

FLOAT32, FLOAT64, COMPLEX64, COMPLEX128 = 0, 1, 2, 3
DTYPE_SIZE = {FLOAT32: 4, FLOAT64: 8, COMPLEX64: 8, COMPLEX128: 16}
DTYPE_NAMES = {"f32": FLOAT32, "f64": FLOAT64, "c64": COMPLEX64, "c128": COMPLEX128,
               "float32": FLOAT32, "float64": FLOAT64, "complex64": COMPLEX64,
               "complex128": COMPLEX128}


class G2GBlock(ctypes.Structure):
    _fields_ = [("row", ctypes.c_int32), ("col", ctypes.c_int32),
                ("stride", ctypes.c_int64), ("data", ctypes.c_void_p)]


class G2GLayout(ctypes.Structure):
    _fields_ = [("n_rows_split", ctypes.c_int32), ("n_cols_split", ctypes.c_int32),
                ("rows_split", ctypes.c_void_p), ("cols_split", ctypes.c_void_p),
                ("owners", ctypes.c_void_p), ("n_ranks", ctypes.c_int32),
                ("n_blocks", ctypes.c_int32), ("blocks", ctypes.c_void_p)]


class G2GPlanInfo(ctypes.Structure):
    _fields_ = [("m", ctypes.c_int64), ("n", ctypes.c_int64), ("n_pieces", ctypes.c_int64),
                ("bytes", ctypes.c_int64), ("remote_bytes", ctypes.c_int64),
                ("n_tiles", ctypes.c_int64), ("tile_bytes", ctypes.c_int64),
                ("n_ranks", ctypes.c_int32), ("dtype", ctypes.c_int32),
                ("trans", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class G2GTuning(ctypes.Structure):
    _fields_ = [("tile_bytes", ctypes.c_int64), ("threads", ctypes.c_int32),
                ("unroll_bytes", ctypes.c_int32), ("nontemporal", ctypes.c_int32),
                ("launch_mode", ctypes.c_int32), ("column_chunks", ctypes.c_int32),
                ("stream_order", ctypes.c_int32), ("row_chunks", ctypes.c_int32), ("tail_split", ctypes.c_int32),
                ("ms", ctypes.c_double)]


class G2GServerStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_int64), ("completed", ctypes.c_int64),
                ("rejected", ctypes.c_int64), ("launches", ctypes.c_int64),
                ("bytes", ctypes.c_int64), ("in_flight_bytes", ctypes.c_int64),
                ("peak_in_flight_bytes", ctypes.c_int64)]


class G2GEdge(ctypes.Structure):
    _fields_ = [("src", ctypes.c_int32), ("dest", ctypes.c_int32), ("weight", ctypes.c_int64)]


class G2GError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


_vp = ctypes.c_void_p
_i = ctypes.c_int
_i64 = ctypes.c_int64
_p64 = ctypes.POINTER(ctypes.c_int64)
_p32 = ctypes.POINTER(ctypes.c_int32)

# name -> (restype, argtypes); every symbol the public header declares
PROTOTYPES = {
    "g2g_last_error": (ctypes.c_char_p, []),
    "g2g_version": (_i, []),
    "g2g_synchronize": (_i, [_vp]),
    "g2g_plan_create": (_i, [ctypes.POINTER(G2GLayout), ctypes.POINTER(G2GLayout), _i, ctypes.c_char,
                             _vp, ctypes.POINTER(_vp)]),
    "g2g_plan_destroy": (_i, [_vp]),
    "g2g_plan_get_info": (_i, [_vp, ctypes.POINTER(G2GPlanInfo)]),
    "g2g_transform": (_i, [_vp, _vp, _vp, _vp]),
    "g2g_transform_staged": (_i, [_vp, _vp, _vp, _vp]),
    "g2g_transform_graph": (_i, [_vp, _vp, _vp, _vp]),
    "g2g_transform_layouts": (_i, [ctypes.POINTER(G2GLayout), ctypes.POINTER(G2GLayout), _i, ctypes.c_char,
                                   _vp, _vp, _vp]),
    "g2g_plan_cache_set_capacity": (_i, [_i64]),
    "g2g_plan_cache_clear": (_i, []),
    "g2g_plan_cache_stats": (_i, [_p64, _p64, _p64]),
    "g2g_plan_counts": (_i, [_vp, _p64, _i64]),
    "g2g_plan_wire": (_i, [_vp, _p32, _p32, _p32, _p32, _p32, _p32, _p64, _p64, _i64]),
    "g2g_plan_send_buffer": (_i, [_vp, _i, ctypes.POINTER(_vp), _p64]),
    "g2g_batch_create": (_i, [_i, ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.POINTER(_vp), _vp,
                              ctypes.POINTER(_vp)]),
    "g2g_batch_execute": (_i, [_vp, _vp]),
    "g2g_batch_execute_graph": (_i, [_vp, _vp]),
    "g2g_batch_bytes": (_i, [_vp, _p64]),
    "g2g_batch_autotune": (_i, [_vp, _vp, _i, ctypes.POINTER(G2GTuning)]),
    "g2g_batch_get_tuning": (_i, [_vp, ctypes.POINTER(G2GTuning)]),
    "g2g_batch_set_tuning": (_i, [_vp, ctypes.POINTER(G2GTuning), _vp]),
    "g2g_batch_destroy": (_i, [_vp]),
    "g2g_transform_plans_noalloc": (_i, [_i, ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.POINTER(_vp), _vp]),
    "g2g_set_tuning": (_i, [_i64, _i]),
    "g2g_plan_get_tuning": (_i, [_vp, ctypes.POINTER(G2GTuning)]),
    "g2g_plan_set_tuning": (_i, [_vp, ctypes.POINTER(G2GTuning)]),
    "g2g_plan_autotune": (_i, [_vp, _vp, _i, ctypes.POINTER(G2GTuning)]),
    "g2g_server_create": (_i, [_i64, _i, _i, _i, ctypes.POINTER(_vp)]),
    "g2g_server_submit": (_i, [_vp, _vp, _vp, _vp, ctypes.POINTER(_i64)]),
    "g2g_server_poll": (_i, [_vp, _i64, _p32, ctypes.POINTER(ctypes.c_double)]),
    "g2g_server_wait": (_i, [_vp, _i64, ctypes.POINTER(ctypes.c_double)]),
    "g2g_server_drain": (_i, [_vp]),
    "g2g_server_get_stats": (_i, [_vp, ctypes.POINTER(G2GServerStats)]),
    "g2g_server_destroy": (_i, [_vp]),
    "g2g_comm_volume": (_i, [ctypes.POINTER(G2GLayout), ctypes.POINTER(G2GLayout), ctypes.POINTER(G2GEdge),
                             _i64, _p64, _vp]),
    "g2g_optimal_reordering": (_i, [ctypes.POINTER(G2GEdge), _i64, ctypes.c_int32, _p32, _p32]),
    "g2g_fill_layout": (_i, [ctypes.POINTER(G2GLayout), _i, ctypes.c_uint64, _vp]),
    "g2g_scale_layout": (_i, [ctypes.POINTER(G2GLayout), _i, _vp, _vp]),
    "g2g_copy_bytes": (_i, [_vp, _vp, _i64]),
    "g2g_stream_probe": (_i, [_i, _vp, _vp, _i64, _i, _vp, ctypes.POINTER(ctypes.c_double)]),
    "g2g_verify_layout": (_i, [ctypes.POINTER(G2GLayout), _i, ctypes.c_uint64, ctypes.c_char, _i64, _vp,
                               _p64]),
}

_lib = None


def load():
    """Load the built library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m grid2grid_amd.build`")
    # torch ships its own libamdhip64.so.7; when torch is present it must be
    # loaded first so that the library binds to that same HIP runtime (one
    # set of streams and queues per process) instead of a second copy.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in PROTOTYPES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc):
    if rc != G2G_OK:
        msg = load().g2g_last_error()
        raise G2GError(rc, msg.decode() if msg else "")
    return rc
