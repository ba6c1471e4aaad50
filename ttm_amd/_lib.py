"""ctypes binding of ``libttm_amd.so`` (the extern "C" surface of include/ttm_amd.h).

Loading fails loudly when the library has not been built: there is no CPU
fallback anywhere in the package.
"""
from __future__ import annotations

import ctypes
import enum
import os
import pathlib

LIB_PATH = pathlib.Path(__file__).resolve().parent / "lib" / "libttm_amd.so"

_size_p = ctypes.POINTER(ctypes.c_size_t)
_vp = ctypes.c_void_p


class Status(enum.IntEnum):
    SUCCESS = 0
    ORDER_ZERO = 1
    MODE = 2
    NULL_A = 3
    NULL_B = 4
    NULL_C = 5
    NULL_NA = 6
    NULL_NB = 7
    NULL_NC = 8
    NULL_WA = 9
    NULL_WC = 10
    NULL_PIA = 11
    CONTRACTION_DIM = 12
    FREE_DIM = 13
    SHAPE_A = 14
    SHAPE_C = 15
    LAYOUT_A = 16
    STRIDES_A = 17
    STRIDES_C = 18
    NULL_PIB = 19
    NOT_DENSE = 20
    SHAPE_C_MISMATCH = 21
    PRECISION = 22
    HIP = 23
    ALLOC = 24
    ARGUMENT = 25
    POLICY = 26


class Precision(enum.IntEnum):
    DEFAULT = 0
    BF16X3 = 1
    BF16 = 2
    FP16 = 3
    FP8_E4M3 = 4
    FP16X2 = 5


# The reference's policy tags (include/tlib/detail/tags.h:22-69) as the host
# executor's enums.
CPU_PARALLEL = {"sequential": 0, "parallel_blas": 1, "parallel_loop": 2, "parallel_taskloop": 3,
                "parallel_task": 4, "parallel_loop_blas": 5, "batched_gemm": 6, "combined": 7}
CPU_SLICING = {"slice": 0, "subtensor": 1, "combined": 2}
CPU_FUSION = {"none": 0, "outer": 1, "all": 2}

PRECISION_NAMES = {
    "default": Precision.DEFAULT, "fp32": Precision.DEFAULT, "fp64": Precision.DEFAULT,
    "bf16x3": Precision.BF16X3, "bf16": Precision.BF16, "fp16": Precision.FP16,
    "fp8": Precision.FP8_E4M3, "fp8_e4m3": Precision.FP8_E4M3, "fp16x2": Precision.FP16X2,
}

# (name, restype, argtypes) for every symbol of include/ttm_amd.h
_TTM_ARGS = [ctypes.c_uint, ctypes.c_uint,
             _vp, _size_p, _size_p, _size_p,
             _vp, _size_p, _size_p,
             _vp, _size_p, _size_p]
SIGNATURES = [
    ("ttm_sttm", ctypes.c_int, _TTM_ARGS + [_vp]),
    ("ttm_dttm", ctypes.c_int, _TTM_ARGS + [_vp]),
    ("ttm_sttm_ex", ctypes.c_int, _TTM_ARGS + [ctypes.c_int, _vp]),
    ("ttm_sttm_ex_noalloc", ctypes.c_int, _TTM_ARGS + [ctypes.c_int, _vp]),
    ("ttm_sttm_host", ctypes.c_int, _TTM_ARGS),
    ("ttm_dttm_host", ctypes.c_int, _TTM_ARGS),
    ("ttm_sttm_host_ex", ctypes.c_int, _TTM_ARGS + [ctypes.c_int]),
    ("ttm_last_error", ctypes.c_char_p, []),
    ("ttm_status_message", ctypes.c_char_p, [ctypes.c_int]),
    ("ttm_is_device_pointer", ctypes.c_int, [_vp]),
    ("ttm_plan_kernel", ctypes.c_char_p,
     [ctypes.c_uint, ctypes.c_uint, ctypes.c_int, _size_p, _size_p, _size_p, _size_p, ctypes.c_int]),
    ("ttm_plan_kernel_symbol", ctypes.c_char_p,
     [ctypes.c_uint, ctypes.c_uint, ctypes.c_int, _size_p, _size_p, _size_p, _size_p, ctypes.c_int]),
    ("ttm_version", ctypes.c_char_p, []),
    ("ttm_force_kernel", ctypes.c_int, [ctypes.c_char_p]),
    ("ttm_device_alloc", ctypes.c_int, [ctypes.c_size_t, ctypes.POINTER(_vp)]),
    ("ttm_device_free", ctypes.c_int, [_vp]),
    ("ttm_device_copy", ctypes.c_int, [_vp, _vp, ctypes.c_size_t]),
    ("ttm_device_zero", ctypes.c_int, [_vp, ctypes.c_size_t]),
    ("ttm_device_synchronize", ctypes.c_int, []),
    ("ttm_sttm_cpu", ctypes.c_int, _TTM_ARGS + [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint,
                                                 ctypes.c_double]),
    ("ttm_dttm_cpu", ctypes.c_int, _TTM_ARGS + [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint,
                                                 ctypes.c_double]),
    ("ttm_stune", ctypes.c_int, _TTM_ARGS + [_vp, ctypes.POINTER(ctypes.c_uint)]),
    ("ttm_dtune", ctypes.c_int, _TTM_ARGS + [_vp, ctypes.POINTER(ctypes.c_uint)]),
    ("ttm_tune_clear", None, []),
    ("ttm_plan_splits", ctypes.c_uint,
     [ctypes.c_uint, ctypes.c_uint, ctypes.c_int, _size_p, _size_p, _size_p, _size_p, ctypes.c_int]),
    ("ttm_last_kernel", ctypes.c_char_p, []),
    ("ttm_last_splits", ctypes.c_uint, []),
    ("ttm_workspace_bytes", ctypes.c_size_t, []),
]

_lib = None


class TTMError(RuntimeError):
    """A failing status of the C ABI; ``str()`` is the library's message."""

    def __init__(self, status: int, message: str):
        super().__init__(message)
        self.status = Status(status) if status in Status._value2member_map_ else status


def lib() -> ctypes.CDLL:
    """The loaded library (loads it on first use)."""
    global _lib
    if _lib is None:
        path = os.environ.get("TTM_AMD_LIB", str(LIB_PATH))
        if not os.path.exists(path):
            raise ImportError(
                f"libttm_amd.so not found at {path}: build it with `python -m ttm_amd.build` "
                "(there is no CPU fallback)")
        handle = ctypes.CDLL(path)
        for name, res, args in SIGNATURES:
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def check(status: int) -> None:
    if status != 0:
        raise TTMError(status, lib().ttm_last_error().decode())


def sizes(values) -> ctypes.Array:
    """A C ``size_t`` array (kept alive by the caller)."""
    vals = [int(v) for v in values]
    return (ctypes.c_size_t * max(1, len(vals)))(*vals)


def force_kernel(name: str | None) -> None:
    """Testing hook: force a kernel family ('' / None = planner's choice)."""
    if lib().ttm_force_kernel((name or "").encode()) != 0:
        raise ValueError(f"unknown kernel family {name!r}")


def plan_kernel(q: int, na, pia, nb, pib, is_double: bool = False, precision: int = 0) -> str:
    """Name of the kernel the planner picks (no GPU needed)."""
    p = len(na)
    return lib().ttm_plan_kernel(q, p, int(is_double), sizes(na), sizes(pia), sizes(nb), sizes(pib),
                                 int(precision)).decode()


def plan_kernel_symbol(q: int, na, pia, nb, pib, is_double: bool = False, precision: int = 0) -> str:
    """Kernel template instance the planner launches (as rocprof names it)."""
    p = len(na)
    return lib().ttm_plan_kernel_symbol(q, p, int(is_double), sizes(na), sizes(pia), sizes(nb), sizes(pib),
                                        int(precision)).decode()


def plan_splits(q: int, na, pia, nb, pib, is_double: bool = False, precision: int = 0) -> int:
    """Contraction splits the planner uses for a dense, aligned problem."""
    p = len(na)
    return int(lib().ttm_plan_splits(q, p, int(is_double), sizes(na), sizes(pia), sizes(nb), sizes(pib),
                                     int(precision)))


def last_kernel() -> tuple[str, int]:
    """(kernel template instance, contraction splits) of this thread's last product."""
    L = lib()
    return L.ttm_last_kernel().decode(), int(L.ttm_last_splits())


def tune_clear() -> None:
    """Forget every tuned contraction split of this process."""
    lib().ttm_tune_clear()


def workspace_bytes() -> int:
    """Device memory held by the current device's workspace pool."""
    return int(lib().ttm_workspace_bytes())


def version() -> str:
    return lib().ttm_version().decode()
