"""``ttmpy``: the reference's Python module (ttmpy/src/wrapped_ttm.cpp) on the MI355X.

Same functions, argument meaning and error behaviour:

* ``ttm(q, A, B)`` -- mode-q product of an order-p array A (C-contiguous,
  i.e. the last-order layout, as the reference assumes at wrapped_ttm.cpp:45)
  with a row-major m x n_q matrix B (:46); returns C with extent m in mode q.
* ``ttms(q, A, Bs, order='forward')`` -- the multi-TTM chain over every mode
  except q, in 'forward', 'backward' or 'optimal' order (:81-222).

Extensions over the reference (which registers float64 NumPy arrays only,
:229-230): float32 is kept as float32, ROCm torch tensors are accepted and
stay on the device, and an optional ``precision`` selects a float32 arithmetic.
Arrays of other dtypes are converted to float64, as pybind11's array_t does.
"""
from __future__ import annotations

import numpy as np

from . import api

try:
    import torch
except ImportError:  # pragma: no cover
    torch = None


def _as_input(x):
    if torch is not None and isinstance(x, torch.Tensor):
        if x.dtype not in (torch.float32, torch.float64):
            x = x.to(torch.float64)
        return x
    x = np.asarray(x)
    if x.dtype not in (np.float32, np.float64):
        x = x.astype(np.float64)
    return x


def ttm(q, A, B, precision=None):
    """C = A x_q B (wrapped_ttm.cpp:23-76)."""
    A = _as_input(A)
    B = _as_input(B)
    p = A.ndim
    if p == 0:
        raise ValueError("Error calling ttmpy::ttm: first input should be a tensor with order greater than zero.")
    if q < 1 or q > p:
        raise ValueError("Error calling ttmpy::ttm: contraction mode should be greater than zero and at most equal to p.")
    if B.ndim != 2:
        raise ValueError("Error calling ttmpy::ttm: second input should be a mtrix with order equal to 2.")
    if B.dtype != A.dtype:
        B = B.to(A.dtype) if torch is not None and isinstance(B, torch.Tensor) else B.astype(A.dtype)
    na = list(A.shape)
    wa = api._elem_strides(A)
    if any(s < 0 for s in wa):
        raise ValueError("Error calling ttmpy::ttm: negative strides are not supported.")
    pia = list(range(p, 0, -1))  # last-order layout
    pib = [2, 1]                  # row-major B
    Bc = B.contiguous() if torch is not None and isinstance(B, torch.Tensor) else np.ascontiguousarray(B)
    nb = list(B.shape)
    nc = list(na)
    nc[q - 1] = nb[0]
    C = api.empty_like_layout(A, nc, pia)
    wc = api._elem_strides(C)
    api.ttm_raw(q, p, A, na, wa, pia, Bc, nb, pib, C, nc, wc, precision=precision)
    return C


def ttms(q, A, Bs, order="forward", precision=None):
    """Multi-TTM over all modes except q (wrapped_ttm.cpp:81-222).

    Bs holds p-1 matrices; Bs[r-1] multiplies mode r for r < q and Bs[r-2]
    mode r for r > q.  'forward' contracts modes in ascending order,
    'backward' descending, 'optimal' ascending in 1/n_k - 1/m_k of each
    matrix B_k (m_k x n_k).
    """
    if order not in ("optimal", "backward", "forward"):
        raise ValueError("Error calling ttmpy::ttms: multiplication order should be either 'optimal', 'backward' or 'forward'.")
    A = _as_input(A)
    p = A.ndim
    if p == 0:
        raise ValueError("Error calling ttmpy::ttms: input tensor order should be greater than zero.")
    Bs = [_as_input(b) for b in Bs]
    if len(Bs) != p - 1:
        raise ValueError("Error calling ttmpy::ttms: number of input matrices is not equal to the tensor order - 1.")
    if q < 1 or q > p:
        raise ValueError("Error calling ttmpy::ttms: non-contraction mode should be greater than zero and at most equal to p.")
    if not all(b.ndim == 2 for b in Bs):
        raise ValueError("Error calling ttmpy::ttms: some of the input matrices are not matrices.")
    na = list(A.shape)
    # mode r (one-based, r != q) is multiplied by Bs[idx(r)]
    idx = {r: (r - 1 if r < q else r - 2) for r in range(1, p + 1) if r != q}
    if any(Bs[i].shape[1] != na[r - 1] for r, i in idx.items()):
        raise ValueError("Error calling ttmpy::ttms: vector dimension is not compatible with the dimension of a tensor mode.")

    if order == "forward":
        modes = [r for r in range(1, p + 1) if r != q]
    elif order == "backward":
        modes = [r for r in range(p, 0, -1) if r != q]
    else:
        # stable ascending sort by beta_k = 1/n_k - 1/m_k over modes 1..q-1, q+1..p
        def beta(r):
            m_k, n_k = Bs[idx[r]].shape
            return 1.0 / float(n_k) - 1.0 / float(m_k)
        modes = sorted((r for r in range(1, p + 1) if r != q), key=beta)

    if _device_chain_ok(A, Bs):
        # NumPy inputs: upload A and the matrices once, keep every
        # intermediate on the device, copy the result back once
        dev = torch.device("cuda")
        C = torch.from_numpy(A).to(dev)
        Bd = [torch.from_numpy(np.ascontiguousarray(b)).to(dev) for b in Bs]
        for r in modes:
            C = ttm(r, C, Bd[idx[r]], precision=precision)
        return C.cpu().numpy()
    C = A
    for r in modes:
        C = ttm(r, C, Bs[idx[r]], precision=precision)
    return C


def _device_chain_ok(A, Bs):
    """Host (NumPy) inputs in the reference's C-contiguous layout, with a GPU
    present: the chain can run device-resident.  Other layouts take the
    per-call path, which reports the reference's stride error."""
    if torch is None or not isinstance(A, np.ndarray) or not A.flags.c_contiguous or A.size == 0:
        return False
    if not all(isinstance(b, np.ndarray) for b in Bs):
        return False
    try:
        return torch.cuda.is_available()
    except Exception:  # pragma: no cover
        return False
