"""Tucker decomposition by truncated HOSVD, on the batched multi-TTM chains.

The caller one step above the chain engine: for each tensor A and ranks
(r_1, ..., r_p),

* factor k = the r_k leading left singular vectors of the mode-k unfolding
  A_(k), taken as the top eigenvectors of the Gram matrix A_(k) A_(k)^T
  (n_k x n_k, float64, on the device).  The Gram matrix is itself a TTM of
  this library: with X = A_(k) as a two-way tensor (a float64 copy, mode k
  slowest), G = X x_2 X contracts X's long mode with the rows of X (the
  split-contraction tile, n_k * N_A multiply-adds per mode);
* core = A x_1 U_1^T x_2 U_2^T ... x_p U_p^T, the multi-TTM chain with factors
  of orthonormal rows -- every core of a batch comes from one chain plan
  (one persistent launch per stage, hipGraph replay);
* ``reconstruct`` multiplies a core back by the factors (again one plan).

The eigendecompositions (n_k x n_k, n_k at most a few hundred) are torch
calls; every product over the tensor is this library's.
"""
from __future__ import annotations

from .chain import ChainPlan


def _gram(A, k):
    """A_(k) A_(k)^T in float64: X = A_(k) (n_k x M, M contiguous) as a
    two-way tensor of extents (n_k, M) with mode 2 fastest, and
    G[i][j] = sum_l X[j][l] X[i][l] = (X x_2 X)[i][j], with X itself as the
    row-major n_k x M matrix B."""
    import torch
    from .api import ttm_raw
    p = A.dim()
    order = [k] + [r for r in range(p) if r != k]
    X = A.permute(*order).reshape(A.shape[k], -1).to(torch.float64).contiguous()
    n, M = X.shape
    G = torch.empty((n, n), dtype=torch.float64, device=X.device)
    ttm_raw(2, 2, X, [n, M], [M, 1], [2, 1], X, [n, M], [2, 1], G, [n, n], [n, 1])
    return G


def factors_of(A, ranks):
    """Factors with orthonormal rows, U_k^T of shape r_k x n_k, in A's dtype."""
    import torch
    for k, r in enumerate(ranks):
        n = A.shape[k]
        if not 1 <= r <= n:
            raise ValueError(f"rank of mode {k + 1} must be in 1..{n}")
    out = []
    for k, r in enumerate(ranks):
        w, V = torch.linalg.eigh(_gram(A, k))           # ascending eigenvalues
        Ut = V[:, torch.argsort(w, descending=True)[:r]].transpose(0, 1)
        out.append(Ut.to(A.dtype).contiguous())
    return out


def _empty_core(torch, A, ranks):
    # the core keeps A's layout (the chain writes C in A's layout)
    from .api import empty_like_layout, layout_from_strides
    return empty_like_layout(A, list(ranks), layout_from_strides(list(A.shape), list(A.stride())))


def hosvd_batch(tensors, ranks_list, precision=None):
    """Truncated HOSVD of many tensors (ROCm torch, float32/float64, any dense
    layout): returns a list of (core, [U_1^T, ..., U_p^T])."""
    import torch
    tensors = list(tensors)
    facs = [factors_of(A, rk) for A, rk in zip(tensors, ranks_list)]
    cores = [_empty_core(torch, A, rk) for A, rk in zip(tensors, ranks_list)]
    plan = ChainPlan(list(zip(tensors, facs, cores)), precision=precision)
    plan.run()
    torch.cuda.current_stream().synchronize()  # the plan's intermediates are freed below
    plan.close()
    return list(zip(cores, facs))


def hosvd(A, ranks, precision=None):
    """Truncated HOSVD of one tensor: (core, [U_1^T, ..., U_p^T])."""
    return hosvd_batch([A], [ranks], precision=precision)[0]


def reconstruct_batch(cores, factor_lists, precision=None):
    """core x_1 U_1 x_2 U_2 ... for every (core, [U_k^T]) of a batch."""
    import torch
    outs, items = [], []
    for G, facs in zip(cores, factor_lists):
        Us = [Ut.transpose(0, 1).contiguous() for Ut in facs]   # n_k x r_k
        C = _empty_core(torch, G, [U.shape[0] for U in Us])
        items.append((G, Us, C))
        outs.append(C)
    plan = ChainPlan(items, precision=precision)
    plan.run()
    torch.cuda.current_stream().synchronize()
    plan.close()
    return outs


def reconstruct(core, factors, precision=None):
    return reconstruct_batch([core], [factors], precision=precision)[0]


__all__ = ["factors_of", "hosvd", "hosvd_batch", "reconstruct", "reconstruct_batch"]
