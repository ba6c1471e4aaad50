"""Python front end of the C ABI for NumPy arrays (host) and ROCm torch tensors (device).

``ttm_raw`` mirrors the reference's C-like low-level call
(include/tlib/ttm.h:54-93) argument for argument; ``ttm`` is a tensor-level
convenience that infers the layout of A from its strides.
"""
from __future__ import annotations

import ctypes

import numpy as np

from ._lib import (CPU_FUSION, CPU_PARALLEL, CPU_SLICING, PRECISION_NAMES, Precision, Status, TTMError, check, lib,
                   sizes)

try:  # torch is optional for host-only use
    import torch
except ImportError:  # pragma: no cover
    torch = None


def _is_torch(x) -> bool:
    return torch is not None and isinstance(x, torch.Tensor)


def _dtype_of(x):
    if _is_torch(x):
        return {torch.float32: np.float32, torch.float64: np.float64}.get(x.dtype)
    return np.dtype(x.dtype).type


def _pointer(x) -> int:
    if _is_torch(x):
        return x.data_ptr()
    if isinstance(x, np.ndarray):
        return x.ctypes.data
    if isinstance(x, int):
        return x
    raise TypeError(f"unsupported buffer type {type(x).__name__}")


def _residency(x) -> str:
    """'device' (ROCm torch tensor), 'host' (NumPy array or CPU torch tensor)
    or 'raw' (an integer device address)."""
    if _is_torch(x):
        return "device" if x.is_cuda else "host"
    if isinstance(x, np.ndarray):
        return "host"
    if isinstance(x, int) and not isinstance(x, bool):
        return "raw"
    raise TypeError(f"unsupported buffer type {type(x).__name__}")


def _dense_numel(x):
    """Elements of x when its strides describe a gap-free, non-overlapping
    block starting at its data pointer (any mode order), else None."""
    if _is_torch(x):
        shape, strides = list(x.shape), list(x.stride())
    else:
        item = x.dtype.itemsize
        if any(s % item for s in x.strides):
            return None
        shape, strides = list(x.shape), [s // item for s in x.strides]
    n = 1
    for e in shape:
        n *= int(e)
    if n == 0:
        return 0
    expect = 1
    for s, e in sorted((s, e) for s, e in zip(strides, shape) if e > 1):
        if s != expect:
            return None
        expect *= e
    return n


def _prod(xs) -> int:
    n = 1
    for v in xs:
        n *= int(v)
    return n


def _check_buffers(p, a, na, b, nb, c, nc, dtype=None):
    """Residency, element type and size of A, B and C, checked before any
    call into the library (which sees only pointers).  Returns (device, dt)."""
    res = {_residency(x) for x in (a, b, c)}
    if len(res) != 1:
        raise TypeError("A, B and C must all be ROCm device tensors, all host arrays, or all raw device "
                        f"addresses (got {sorted(res)})")
    kind = res.pop()
    if kind == "raw":
        if dtype is None:
            raise TypeError("raw device addresses need dtype= (float32 or float64)")
        dt = np.dtype(dtype).type
        if dt not in (np.float32, np.float64):
            raise TypeError("ttm supports float32 and float64 elements")
        return True, dt
    dt = _dtype_of(a)
    if dt not in (np.float32, np.float64):
        raise TypeError("ttm supports float32 and float64 elements")
    for name, x in (("B", b), ("C", c)):
        if _dtype_of(x) != dt:
            raise TypeError(f"{name} must have A's element type {np.dtype(dt).name}")
    if kind == "device" and not (a.device == b.device == c.device):
        raise TypeError("A, B and C must be on the same device")
    # Buffer sizes, when the extents are well formed (malformed metadata is
    # left to the library, which reports the reference's error for it).
    try:
        shapes_ok = (len(na) == p and len(nc) == p and len(nb) == 2 and
                     all(int(v) >= 0 for v in (*na, *nb, *nc)))
    except TypeError:
        shapes_ok = False
    if shapes_ok:
        for name, x, need in (("A", a, _prod(na)), ("B", b, _prod(nb)), ("C", c, _prod(nc))):
            have = _dense_numel(x)
            if have is None:
                raise ValueError(f"{name} must be one dense block of memory")
            if have < need:
                raise ValueError(f"{name} holds {have} elements, its extents need {need}")
    return kind == "device", dt


def _precision(precision) -> int:
    if isinstance(precision, str):
        try:
            return int(PRECISION_NAMES[precision.lower()])
        except KeyError:
            raise ValueError(f"unknown precision {precision!r}") from None
    return int(Precision(precision if precision is not None else 0))


def _stream_handle(stream) -> int:
    if stream is not None:
        return int(getattr(stream, "cuda_stream", stream))
    return int(torch.cuda.current_stream().cuda_stream)


def ttm_raw(q, p, a, na, wa, pia, b, nb, pib, c, nc, wc, *, precision=None, stream=None, dtype=None):
    """The reference's low-level call: C (dense, layout pia) = A x_q B.

    a, b, c are NumPy arrays (host memory, synchronous call), torch tensors on
    a ROCm device (asynchronous on ``stream``, default the current stream), or
    raw device addresses (all three, with ``dtype``).  Mixed residencies,
    element types other than A's and buffers smaller than their extents raise
    before the library is called.  na, wa, pia, nb, pib, nc, wc are integer sequences
    as in the reference.  Raises ``TTMError`` with the reference's message on
    invalid arguments.
    """
    device, dt = _check_buffers(p, a, na, b, nb, c, nc, dtype)
    prec = _precision(precision)
    args = [ctypes.c_uint(int(q)), ctypes.c_uint(int(p)),
            _pointer(a), sizes(na), sizes(wa), sizes(pia),
            _pointer(b), sizes(nb), sizes(pib),
            _pointer(c), sizes(nc), sizes(wc)]
    L = lib()
    if dt == np.float64 and prec:
        raise TTMError(int(Status.PRECISION), L.ttm_status_message(int(Status.PRECISION)).decode())
    if device:
        s = ctypes.c_void_p(_stream_handle(stream))
        if dt == np.float32:
            st = L.ttm_sttm_ex(*args, prec, s) if prec else L.ttm_sttm(*args, s)
        else:
            st = L.ttm_dttm(*args, s)
    elif dt == np.float32:
        st = L.ttm_sttm_host_ex(*args, prec) if prec else L.ttm_sttm_host(*args)
    else:
        st = L.ttm_dttm_host(*args)
    check(st)
    return c


def ttm_cpu_raw(q, p, a, na, wa, pia, b, nb, pib, c, nc, wc, *, parallel="combined", slicing="combined",
                fusion="all", threads=0, ratio=0.0):
    """The reference's low-level call on the host executor (no GPU): NumPy
    arrays (or CPU torch tensors), the policy tag triple by name
    (parallel: sequential, parallel_blas, parallel_loop, parallel_taskloop,
    parallel_task, parallel_loop_blas, batched_gemm, combined; slicing:
    slice, subtensor, combined; fusion: none, outer, all).  A triple the
    reference does not define raises TTMError with status POLICY."""
    device, dt = _check_buffers(p, a, na, b, nb, c, nc)
    if device:
        raise TypeError("the host executor takes host arrays")
    L = lib()
    args = [ctypes.c_uint(int(q)), ctypes.c_uint(int(p)),
            _pointer(a), sizes(na), sizes(wa), sizes(pia),
            _pointer(b), sizes(nb), sizes(pib),
            _pointer(c), sizes(nc), sizes(wc),
            CPU_PARALLEL[parallel], CPU_SLICING[slicing], CPU_FUSION[fusion], int(threads), float(ratio)]
    check((L.ttm_sttm_cpu if dt == np.float32 else L.ttm_dttm_cpu)(*args))
    return c


def tune_raw(q, p, a, na, wa, pia, b, nb, pib, c, nc, wc, *, stream=None, dtype=None):
    """Time the contraction-split counts of this product on its own device
    buffers (C receives the product) and keep the fastest for its shape class
    for later calls of the process.  Returns the chosen split count (1 for
    kernels that do not split)."""
    device, dt = _check_buffers(p, a, na, b, nb, c, nc, dtype)
    if not device:
        raise TypeError("tuning needs device buffers (ROCm torch tensors or raw device addresses)")
    L = lib()
    out = ctypes.c_uint(0)
    args = [ctypes.c_uint(int(q)), ctypes.c_uint(int(p)),
            _pointer(a), sizes(na), sizes(wa), sizes(pia),
            _pointer(b), sizes(nb), sizes(pib),
            _pointer(c), sizes(nc), sizes(wc), ctypes.c_void_p(_stream_handle(stream)), ctypes.byref(out)]
    check((L.ttm_stune if dt == np.float32 else L.ttm_dtune)(*args))
    return int(out.value)


def layout_from_strides(shape, strides):
    """One-based layout (fastest mode first) of a dense tensor from its element
    strides; extent-1 modes are placed by stride, ties by mode index."""
    order = sorted(range(len(shape)), key=lambda r: (strides[r], r))
    return [r + 1 for r in order]


def dense_strides(shape, layout):
    """Dense element strides of ``shape`` in ``layout`` (no vector special case)."""
    w = [0] * len(shape)
    s = 1
    for mode in layout:
        w[mode - 1] = s
        s *= shape[mode - 1]
    return w


def _elem_strides(x):
    if _is_torch(x):
        return list(x.stride())
    item = x.dtype.itemsize
    if any(s % item for s in x.strides):
        raise ValueError("strides must be multiples of the element size")
    return [s // item for s in x.strides]


def empty_like_layout(template, shape, layout):
    """Uninitialised dense tensor of ``shape`` in ``layout`` on ``template``'s device."""
    w = dense_strides(shape, layout)
    if _is_torch(template):
        return torch.empty_strided(tuple(shape), tuple(w), dtype=template.dtype, device=template.device)
    dt = np.dtype(template.dtype)
    buf = np.empty(int(np.prod(shape)) if len(shape) else 1, dtype=dt)
    return np.lib.stride_tricks.as_strided(buf, shape=tuple(shape), strides=tuple(s * dt.itemsize for s in w))


def ttm(q, A, B, *, layout=None, b_layout="row", precision=None, out=None, stream=None):
    """C = A x_q B for a dense A in any layout.

    ``layout`` (one-based, fastest mode first) defaults to the one A's strides
    describe.  ``B`` is m x n_q, read as row-major (``b_layout='row'``, NumPy's
    default) or column-major ('col'); it is made contiguous in that order.
    C is returned in A's layout (or written to ``out``).
    """
    p = A.ndim
    if p == 0:
        raise ValueError("ttm: A must have order greater than zero")
    if not 1 <= q <= p:
        raise ValueError("ttm: contraction mode must satisfy 1 <= q <= A.ndim")
    if B.ndim != 2:
        raise ValueError("ttm: B must be a matrix")
    na = list(A.shape)
    wa = _elem_strides(A)
    if any(s < 0 for s in wa):
        raise ValueError("ttm: negative strides are not supported")
    pia = list(layout) if layout is not None else layout_from_strides(na, wa)
    colmajor = b_layout in ("col", "column", "F")
    # B takes A's element type and residency; `out` must already have both.
    if _is_torch(A):
        if not _is_torch(B):
            B = torch.from_numpy(np.asarray(B))
        B = B.to(device=A.device, dtype=A.dtype)
    else:
        B = B.cpu().numpy() if _is_torch(B) else np.asarray(B)
        B = B.astype(A.dtype, copy=False)
    if out is not None:
        if _residency(out) != _residency(A) or _dtype_of(out) != _dtype_of(A):
            raise TypeError("out must have A's element type and residency")
    if _is_torch(B):
        Bc = B.t().contiguous().t() if colmajor else B.contiguous()
    else:
        Bc = np.asfortranarray(B) if colmajor else np.ascontiguousarray(B)
    nb = list(B.shape)
    pib = [1, 2] if colmajor else [2, 1]
    nc = list(na)
    nc[q - 1] = nb[0]
    C = out if out is not None else empty_like_layout(A, nc, pia)
    wc = _elem_strides(C)
    ttm_raw(q, p, A, na, wa, pia, Bc, nb, pib, C, nc, wc, precision=precision, stream=stream)
    return C
