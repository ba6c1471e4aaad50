"""Batched multi-TTM chains on the GPU (ttm_chain_* of include/ttm_amd.h).

``ChainPlan(items)`` prepares one plan for many tensors of uneven extents; each
item is ``(A, factors, C)`` with ROCm torch tensors: A and C dense in the same
layout, ``factors[k]`` the row-major ``r_k x n_k`` factor of mode k+1 or None.
A mode whose factor is None keeps its extent: its stage only moves the data
into the next stage's layout (a batched transpose, one read and one write of
the tensor), no product.
``plan.run()`` launches one persistent grouped kernel per stage,
``plan.run_graph()`` replays the whole chain as one hipGraph, ``plan.tune()``
picks tile shapes by timing them.
"""
from __future__ import annotations

import ctypes

from ._lib import Status, TTMError, check, lib
from .api import _precision, _stream_handle, dense_strides, empty_like_layout, layout_from_strides

MAX_ORDER = 8


class ChainDesc(ctypes.Structure):
    _fields_ = [
        ("p", ctypes.c_uint),
        ("n", ctypes.c_size_t * MAX_ORDER),
        ("pi", ctypes.c_size_t * MAX_ORDER),
        ("a", ctypes.c_void_p),
        ("u", ctypes.c_void_p * MAX_ORDER),
        ("r", ctypes.c_size_t * MAX_ORDER),
        ("c", ctypes.c_void_p),
    ]


_bound = False


def _bind():
    global _bound
    if _bound:
        return lib()
    L = lib()
    P = ctypes.c_void_p
    L.ttm_chain_plan_create.restype = ctypes.c_int
    L.ttm_chain_plan_create.argtypes = [ctypes.c_size_t, ctypes.POINTER(ChainDesc), ctypes.c_int, ctypes.c_int,
                                        ctypes.POINTER(P)]
    for name in ("ttm_chain_run", "ttm_chain_run_graph", "ttm_chain_tune"):
        getattr(L, name).restype = ctypes.c_int
        getattr(L, name).argtypes = [P, P]
    L.ttm_chain_stages.restype = ctypes.c_uint
    L.ttm_chain_stages.argtypes = [P]
    L.ttm_chain_stage_kernel.restype = ctypes.c_char_p
    L.ttm_chain_stage_kernel.argtypes = [P, ctypes.c_uint]
    L.ttm_chain_plan_destroy.restype = None
    L.ttm_chain_plan_destroy.argtypes = [P]
    _bound = True
    return L


def _dense_in(shape, strides, layout):
    """Dense in `layout`, ignoring the (free) strides of extent-1 modes."""
    want = dense_strides(shape, layout)
    return all(n == 1 or s == w for n, s, w in zip(shape, strides, want))


def make_desc(A, factors, C, layout=None):
    """Descriptor for one tensor (torch tensors on the device).  `layout`
    (one-based, fastest first) defaults to the one A's strides describe; pass
    it when A has extent-1 modes whose place in C's layout matters."""
    import torch
    p = A.dim()
    if not 1 <= p <= MAX_ORDER:
        raise ValueError(f"chain tensors must have order 1..{MAX_ORDER}")
    if len(factors) != p:
        raise ValueError("one factor (or None) per mode is required")
    if A.dtype != C.dtype or A.dtype not in (torch.float32, torch.float64):
        raise TypeError("A and C must both be float32 or both float64")
    if not (A.is_cuda and C.is_cuda) or A.device != C.device:
        raise TypeError("A and C must be ROCm device tensors on one device")
    na = list(A.shape)
    pia = list(layout) if layout is not None else layout_from_strides(na, list(A.stride()))
    if not _dense_in(na, list(A.stride()), pia):
        raise ValueError("A must be dense in its layout")
    d = ChainDesc()
    d.p = p
    for k in range(p):
        d.n[k] = na[k]
        d.pi[k] = pia[k]
        U = factors[k]
        if U is not None:
            if not isinstance(U, torch.Tensor) or not U.is_cuda or U.device != A.device:
                raise TypeError(f"factor {k + 1} must be a ROCm device tensor on A's device")
            if U.dim() != 2 or U.shape[1] != na[k] or U.dtype != A.dtype or not U.is_contiguous():
                raise ValueError(f"factor {k + 1} must be a contiguous r x {na[k]} matrix of A's dtype")
            d.u[k] = U.data_ptr()
            d.r[k] = U.shape[0]
        else:
            d.u[k] = None
            d.r[k] = na[k]
    nc = [int(d.r[k]) for k in range(p)]
    if list(C.shape) != nc or not _dense_in(nc, list(C.stride()), pia):
        raise ValueError("C must have the chain's output extents, dense in A's layout")
    d.a = A.data_ptr()
    d.c = C.data_ptr()
    return d


class ChainPlan:
    """A plan over many (A, factors, C) or (A, factors, C, layout) items;
    keeps references to the tensors."""

    def __init__(self, items, precision=None):
        import torch
        L = _bind()
        self._items = list(items)
        if not self._items:
            raise ValueError("a chain plan needs at least one tensor")
        dtypes = {it[0].dtype for it in self._items}
        if len(dtypes) != 1:
            raise TypeError("every tensor of a chain plan must have the same element type")
        descs = (ChainDesc * len(self._items))(*[make_desc(*it) for it in self._items])
        self._descs = descs
        is_double = self._items[0][0].dtype == torch.float64
        handle = ctypes.c_void_p()
        check(L.ttm_chain_plan_create(len(self._items), descs, int(is_double), _precision(precision),
                                      ctypes.byref(handle)))
        self._h = handle

    def run(self, stream=None):
        check(_bind().ttm_chain_run(self._h, ctypes.c_void_p(_stream_handle(stream))))

    def run_graph(self, stream=None):
        check(_bind().ttm_chain_run_graph(self._h, ctypes.c_void_p(_stream_handle(stream))))

    def tune(self, stream=None):
        check(_bind().ttm_chain_tune(self._h, ctypes.c_void_p(_stream_handle(stream))))

    @property
    def stage_kernels(self):
        L = _bind()
        return [L.ttm_chain_stage_kernel(self._h, s).decode() for s in range(L.ttm_chain_stages(self._h))]

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _bind().ttm_chain_plan_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:  # pragma: no cover
            pass


def multi_ttm(tensors, factor_lists, precision=None):
    """C_t = A_t x_1 U_t1 x_2 U_t2 ... for every t in one plan run (allocates the C_t)."""
    outs = []
    for A, Us in zip(tensors, factor_lists):
        na = list(A.shape)
        pia = layout_from_strides(na, list(A.stride()))
        nc = [U.shape[0] if U is not None else n for U, n in zip(Us, na)]
        outs.append(empty_like_layout(A, nc, pia))
    import torch
    plan = ChainPlan(list(zip(tensors, factor_lists, outs)), precision=precision)
    plan.run()
    torch.cuda.current_stream().synchronize()  # the plan's intermediates are freed below
    plan.close()
    return outs


__all__ = ["ChainDesc", "ChainPlan", "make_desc", "multi_ttm", "Status", "TTMError"]
