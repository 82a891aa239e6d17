"""ctypes binding of the HIP library (include/exmarg_hip.h).

The library is loaded from exmarg_amd/_lib/libexmarg_hip.so (built in-tree by
exmarg_amd.build).  A missing library or a missing device raises: there is
no other implementation to fall back to.
"""

import ctypes
import os
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# EXMARG_HIP_LIB names another build of the same library (for instance one compiled with other tuning)
LIB_PATH = os.environ.get("EXMARG_HIP_LIB") or os.path.join(HERE, "_lib", "libexmarg_hip.so")

EXM_OK, EXM_ERR_ARG, EXM_ERR_HIP, EXM_ERR_NOMEM, EXM_ERR_HANDLE, EXM_ERR_STATE = 0, -1, -2, -3, -4, -5

# every exported symbol and its (restype, argtypes)
P = ctypes.c_void_p
I32, I64, U64, F32, F64, SZ = ctypes.c_int, ctypes.c_int64, ctypes.c_uint64, ctypes.c_float, ctypes.c_double, ctypes.c_size_t
H = ctypes.c_uint64
HP = ctypes.POINTER(ctypes.c_uint64)
SIGNATURES = {
    "exm_init": [I32],
    "exm_last_error": [ctypes.c_char_p, SZ],
    "exm_device_info": [ctypes.c_char_p, SZ, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int)],
    "exm_free": [H],
    "exm_mem_in_use": [ctypes.POINTER(ctypes.c_int64)],
    "exm_mem_info": [ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)],
    "exm_store_create": [I32, I32, I32, I32, I32, P, F64, F64, F64, F64, HP],
    "exm_store_set_table": [H, I32, P],
    "exm_store_load_files": [H, I32, P, P, P, I32, ctypes.POINTER(ctypes.c_double)],
    "exm_store_fill_synthetic": [H, I32, U64, I64, I32],
    "exm_store_read_table": [H, I32, P],
    "exm_bank_create": [P, I32, I32, I32, I64, HP],
    "exm_bank_taylorf2": [P, P, P, I32, F64, I32, I32, I64, P, F64, P, HP],
    "exm_bank_read": [H, P],
    "exm_filter": [H, P, P, P, P, HP],
    "exm_traj_create": [I32, I64, P, P, P, P, HP],
    "exm_traj_read": [H, P, P, P, P],
    "exm_select_times": [H, I32, I32, I32, I32, P, F32, HP],
    "exm_select_times_seeded": [H, I32, I32, I32, I32, P, F32, HP],
    "exm_draw_integers": [P, I64, I64, P],
    "exm_times_read": [H, P, P],
    "exm_gather": [H, H, P, I64, P, P, I32, H, P],
    "exm_plan_create": [H, I32, I64, P, P, P, HP],
    "exm_plan_gather": [H, H, H, H, P],
    "exm_plan_gather_noalloc": [H, H, H, H, P],
    "exm_distance_weights": [P, I32, I32, I32, F64, P, P, I32, P],
    "exm_plan_set_weights": [H, P, P],
    "exm_plan_gather_lse": [H, H, H, H, P, P],
    "exm_batch_create": [H, I64, I32, HP],
    "exm_batch_add": [H, H, H, P, P, P, P, I32, I32, I32, I32, P, F32, ctypes.POINTER(ctypes.c_int)],
    "exm_batch_add_seeded": [H, H, H, P, P, P, P, I32, I32, I32, I32, P, F32, ctypes.POINTER(ctypes.c_int)],
    "exm_batch_add_seeded_many": [H, H, H, I32, P, P, P, P, I32, I32, I32, I32, P, F32, ctypes.POINTER(ctypes.c_int)],
    "exm_batch_run": [H, P],
    "exm_batch_launch": [H],
    "exm_reserve": [I32, I64, I32, I32],
    "exm_batch_poll": [H, ctypes.POINTER(ctypes.c_int)],
    "exm_batch_collect": [H, P],
    "exm_batch_collect_lse": [H, P, P],
    "exm_batch_reset": [H],
    "exm_batch_info": [H, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int)],
    "exm_set_gather_kernel": [I32],
    "exm_set_gather_piece": [I32],
    "exm_set_gather_precision": [I32],
    "exm_set_gather_tile": [I32],
    "exm_set_filter_workspace": [I64],
    "exm_set_filter_precision": [I32],
    "exm_set_filter_transform": [I32],
    "exm_get_gather_tuning": [ctypes.POINTER(ctypes.c_int)] * 4,
    "exm_profile_enable": [I32],
    "exm_profile_reset": [],
    "exm_profile_get": [ctypes.c_char_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int64)],
    "exm_copy_calibration": [I64, ctypes.POINTER(ctypes.c_double)],
    "exm_debug_fail_alloc": [I64],
    "exm_tables_sum_orders": [P, P, I32, I32, I32, I32, P, P, I32, P, I32, P],
    "exm_source_frame": [P, P, I32, I32, P, P, P, I32, P, I32, P, P, P],
    "exm_source_grid_create": [P, I32, I32, P, P, P, I32, P, I32, P, HP],
    "exm_source_grid_run": [H, P, P, P],
    "exm_source_grid_run_many": [H, I32, P, P, P],
    "exm_strain_segments": [P, I32, I64, F64, I32, I32, P],
    "exm_strain_psd": [P, I32, I64, F64, I32, I32, F64, P],
}


class HipError(RuntimeError):
    pass


_lib = None
_lock = threading.Lock()
_device = None


def load():
    """the loaded library (ctypes.CDLL) with argument types set"""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise ImportError("exmarg HIP library not built: %s is missing (python -m exmarg_amd.build)" % LIB_PATH)
            L = ctypes.CDLL(LIB_PATH)
            for name, args in SIGNATURES.items():
                fn = getattr(L, name)
                fn.restype = ctypes.c_int
                fn.argtypes = args
            _lib = L
    return _lib


def last_error():
    buf = ctypes.create_string_buffer(1024)
    load().exm_last_error(buf, 1024)
    return buf.value.decode(errors="replace")


def check(rc):
    if rc == EXM_OK:
        return
    msg = last_error()
    if rc in (EXM_ERR_ARG, EXM_ERR_HANDLE):
        raise ValueError(msg)
    if rc == EXM_ERR_NOMEM:
        raise MemoryError(msg)
    raise HipError(msg)


def call(name, *args):
    check(getattr(load(), name)(*args))


def init(device=0):
    global _device
    if _device is None:
        call("exm_init", int(device))
        _device = int(device)
    return _device


def device_info():
    init()
    name = ctypes.create_string_buffer(256)
    mem = ctypes.c_int64()
    cu = ctypes.c_int()
    call("exm_device_info", name, 256, ctypes.byref(mem), ctypes.byref(cu))
    return dict(name=name.value.decode(), total_mem=mem.value, num_cu=cu.value)


def mem_in_use():
    v = ctypes.c_int64()
    call("exm_mem_in_use", ctypes.byref(v))
    return v.value


def mem_info():
    f, t = ctypes.c_int64(), ctypes.c_int64()
    call("exm_mem_info", ctypes.byref(f), ctypes.byref(t))
    return f.value, t.value


def profile(on=True):
    call("exm_profile_enable", 1 if on else 0)
    call("exm_profile_reset")


def profile_get(name):
    ms = ctypes.c_double()
    n = ctypes.c_int64()
    call("exm_profile_get", name.encode(), ctypes.byref(ms), ctypes.byref(n))
    return ms.value, n.value


def ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def contiguous(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


class Handle:
    """owner of one library object; released by free() or when collected"""

    def __init__(self, value, kind):
        self.value = int(value)
        self.kind = kind

    def free(self):
        if self.value:
            v, self.value = self.value, 0
            if _lib is not None:
                _lib.exm_free(v)

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def __int__(self):
        return self.value


def new_handle(name, *args, kind=""):
    h = ctypes.c_uint64(0)
    call(name, *args, ctypes.byref(h))
    return Handle(h.value, kind)
