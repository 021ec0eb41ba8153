"""ctypes binding of ``libqtensor_amd.so`` (the C ABI in ``include/qtensor_amd.h``).

The library is built in-tree by ``qtensor_amd/build.py``.  Loading fails loudly
when it is missing: there is no CPU fallback for any engine entry point.
"""
import ctypes
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libqtensor_amd.so")

QB_OK = 0
QB_ERR_INVALID = 1
QB_ERR_HIP = 2
QB_ERR_UNSUPPORTED = 3
QB_ERR_NOT_INIT = 4
QB_C64 = 0
QB_C128 = 1
QB_MAX_STEP_INPUTS = 16
QB_MAX_RANK = 31
QB_ITEM_WORDS = 4

EXPORTED = ("qb_version", "qb_init", "qb_shutdown", "qb_last_error", "qb_step",
            "qb_exec_slots", "qb_execute", "qb_wide_execute", "qb_reduce_terms",
            "qb_greedy_order", "qb_elimination_widths", "qb_rgreedy_order", "qb_place_intervals",
            "qb_plan_create", "qb_plan_execute", "qb_plan_destroy", "qb_check_program",
            "qb_execute_noalloc", "qb_lightcone_classes", "qb_compile_lightcone")
# bytes of the static device workspace behind qb_execute_noalloc
QB_NOALLOC_SCRATCH_BYTES = 64 << 20


class StepDesc(ctypes.Structure):
    _fields_ = [("n_inputs", ctypes.c_int32), ("rank_u", ctypes.c_int32),
                ("sum", ctypes.c_int32), ("dtype", ctypes.c_int32),
                ("masks", ctypes.c_uint32 * QB_MAX_STEP_INPUTS)]


class ExecDesc(ctypes.Structure):
    _fields_ = [("program", ctypes.c_void_p), ("items", ctypes.c_void_p),
                ("n_items", ctypes.c_int32), ("item_threads", ctypes.c_int32),
                ("lds_elems", ctypes.c_int32), ("dtype", ctypes.c_int32),
                ("params", ctypes.c_void_p), ("const_pool", ctypes.c_void_p),
                ("scratch", ctypes.c_void_p), ("scratch_elems", ctypes.c_int64),
                ("n_slots", ctypes.c_int32), ("program_words", ctypes.c_int32),
                ("item_out", ctypes.c_void_p), ("queue", ctypes.c_void_p), ("group", ctypes.c_int32)]


class WideDesc(ctypes.Structure):
    _fields_ = [("plan_host", ctypes.c_void_p), ("plan_words", ctypes.c_int64),
                ("plan_dev", ctypes.c_void_p), ("params", ctypes.c_void_p),
                ("arena", ctypes.c_void_p), ("arena_elems", ctypes.c_int64),
                ("slice_bits", ctypes.c_uint32), ("dtype", ctypes.c_int32),
                ("out", ctypes.c_void_p), ("out_stride", ctypes.c_int64), ("n_slices", ctypes.c_int32)]


class PlanDesc(ctypes.Structure):
    _fields_ = [("program", ctypes.c_void_p), ("program_words", ctypes.c_int64),
                ("items", ctypes.c_void_p), ("n_items", ctypes.c_int32), ("n_terms", ctypes.c_int32),
                ("seg", ctypes.c_void_p), ("n_params", ctypes.c_int32), ("item_threads", ctypes.c_int32),
                ("lds_elems", ctypes.c_int32), ("dtype", ctypes.c_int32), ("scratch_elems", ctypes.c_int64),
                ("group", ctypes.c_int32)]


class QBError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"qtensor_amd error {code}: {msg}")
        self.code = code


_lib = None
_lock = threading.Lock()


def load():
    """Load the shared library (no device work)."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing; build it with `python -m qtensor_amd.build` "
                "(the engine has no CPU fallback)")
        # torch carries its own HIP and HSA runtimes; loading it first makes the
        # library bind to those same runtimes (one runtime per process).  Opened
        # first, the library would bring a second pair, and whichever of the two
        # initialises the device second finds none.
        import torch  # noqa: F401
        lib = ctypes.CDLL(LIB_PATH)
        lib.qb_version.restype = ctypes.c_int
        lib.qb_init.argtypes = [ctypes.c_int]
        lib.qb_last_error.restype = ctypes.c_char_p
        lib.qb_step.argtypes = [ctypes.POINTER(StepDesc), ctypes.POINTER(ctypes.c_void_p),
                                ctypes.c_void_p, ctypes.c_void_p]
        lib.qb_exec_slots.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                      ctypes.POINTER(ctypes.c_int32)]
        lib.qb_execute.argtypes = [ctypes.POINTER(ExecDesc), ctypes.c_void_p]
        lib.qb_execute_noalloc.argtypes = [ctypes.POINTER(ExecDesc), ctypes.c_void_p]
        lib.qb_plan_create.argtypes = [ctypes.POINTER(PlanDesc), ctypes.POINTER(ctypes.c_void_p)]
        lib.qb_plan_execute.argtypes = [ctypes.c_void_p] * 5
        lib.qb_plan_destroy.argtypes = [ctypes.c_void_p]
        i32p = ctypes.POINTER(ctypes.c_int32)
        lib.qb_greedy_order.argtypes = [ctypes.c_int32, i32p, i32p, i32p, i32p]
        lib.qb_elimination_widths.argtypes = [ctypes.c_int32, i32p, i32p, ctypes.c_int32, i32p, i32p]
        lib.qb_rgreedy_order.argtypes = [ctypes.c_int32, i32p, i32p, ctypes.c_double, ctypes.c_uint64,
                                         ctypes.c_int32, i32p, i32p]
        i64p = ctypes.POINTER(ctypes.c_int64)
        lib.qb_place_intervals.argtypes = [ctypes.c_int32, i64p, i64p, i64p, i64p, ctypes.c_int64, i64p, i64p]
        lib.qb_wide_execute.argtypes = [ctypes.POINTER(WideDesc), ctypes.c_void_p]
        lib.qb_check_program.argtypes = [i32p, ctypes.c_int64, i32p, ctypes.c_int32, ctypes.c_int32,
                                         ctypes.c_int32, ctypes.c_int32, ctypes.c_int64, ctypes.c_int32]
        lib.qb_reduce_terms.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32,
                                        ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        _lib = lib
        return lib


def check_program(program, items, n_params, item_threads, lds_elems, scratch_elems, group=1):
    """Host check of a batched program and its items (``qb_check_program``)
    before ``qb_execute`` reads them on the device; raises ``ValueError``.
    ``group``: items per executor block (consecutive items sharing a plan)."""
    import numpy as np
    lib = load()
    program = np.ascontiguousarray(program, dtype=np.int32)
    items = np.ascontiguousarray(items, dtype=np.int32).reshape(-1)
    i32p = ctypes.POINTER(ctypes.c_int32)
    check(lib.qb_check_program(program.ctypes.data_as(i32p), program.size,
                               items.ctypes.data_as(i32p) if items.size else ctypes.cast(program.ctypes.data, i32p),
                               items.size // QB_ITEM_WORDS, int(n_params), int(item_threads), int(lds_elems),
                               int(scratch_elems), int(group)))


def check(code):
    if code != QB_OK:
        msg = load().qb_last_error().decode(errors="replace")
        if code == QB_ERR_INVALID:
            raise ValueError(f"qtensor_amd: {msg}")
        raise QBError(code, msg)
    return code


_inited = set()


def init(device=0):
    lib = load()
    if device not in _inited:
        check(lib.qb_init(int(device)))
        _inited.add(device)
    return lib
