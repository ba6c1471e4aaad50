"""ttm_amd -- mode-q tensor-times-matrix product on AMD MI355X (gfx950).

The compute path is ``lib/libttm_amd.so`` (hand-written HIP kernels behind the
extern "C" ABI of ``include/ttm_amd.h``).  This package binds it for NumPy
arrays and ROCm torch tensors:

* :func:`ttm_raw` -- the reference's C-like low-level call;
* :func:`ttm`     -- tensor form, any dense layout;
* :mod:`ttm_amd.ttmpy` -- the reference's ``ttmpy`` module (``ttm``, ``ttms``).
"""
from ._lib import (LIB_PATH, Precision, Status, TTMError, force_kernel, last_kernel, lib, plan_kernel,
                   plan_kernel_symbol, plan_splits, tune_clear, version, workspace_bytes)
from .api import dense_strides, layout_from_strides, ttm, ttm_cpu_raw, ttm_raw, tune_raw
from . import ttmpy

__all__ = ["LIB_PATH", "Precision", "Status", "TTMError", "force_kernel", "last_kernel", "lib", "plan_kernel",
           "plan_kernel_symbol", "plan_splits", "tune_clear", "tune_raw", "version", "workspace_bytes",
           "dense_strides", "layout_from_strides", "ttm", "ttm_cpu_raw", "ttm_raw", "ttmpy"]
from . import chain  # noqa: E402  (batched multi-TTM chains)

__all__.append("chain")
