"""Truncated HOSVD (ttm_amd.tucker) on exactly low-rank tensors.

A = G x_1 U_1 x_2 U_2 x_3 U_3 with orthonormal U_k (built in float64 with
numpy.einsum, the textbook definition) has multilinear rank (r_1, r_2, r_3),
so HOSVD at those ranks reproduces A to rounding: the bar is
||reconstruct(hosvd(A)) - A|| <= 1e-10 ||A|| in float64 and 1e-4 ||A|| in
float32 (also with the FP32-level reduced-precision chains BF16X3 and
FP16X2), and the core's norm equals ||A|| (orthonormal factors).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def low_rank(rng, n, r):
    G = rng.standard_normal(r)
    Us = [np.linalg.qr(rng.standard_normal((nk, rk)))[0] for nk, rk in zip(n, r)]
    A = np.einsum("abc,ia,jb,kc->ijk", G, *Us)
    return A


def to_layout(torch, dev, X, pi, dtype):
    """Device tensor with the values of X (natural order) stored dense in layout pi."""
    order = [m - 1 for m in reversed(pi)]               # slowest mode first
    T = torch.from_numpy(np.ascontiguousarray(np.transpose(X, order))).to(dev, dtype)
    inv = np.argsort(order)
    return T.permute(*[int(i) for i in inv])


@pytest.mark.parametrize("dt", ["f64", "f32", "bf16x3", "fp16x2"])
def test_hosvd_recovers_low_rank_tensor(gpu, dt):
    import torch
    from ttm_amd.tucker import hosvd, reconstruct
    dtype = torch.float64 if dt == "f64" else torch.float32
    prec = dt if dt in ("bf16x3", "fp16x2") else None
    rng = np.random.default_rng(11)
    n, r = (40, 33, 25), (6, 9, 4)
    A = low_rank(rng, n, r)
    At = to_layout(torch, gpu, A, [1, 2, 3], dtype)
    core, facs = hosvd(At, r, precision=prec)
    assert tuple(core.shape) == r and [tuple(f.shape) for f in facs] == [(6, 40), (9, 33), (4, 25)]
    R = reconstruct(core, facs, precision=prec).double().cpu().numpy()
    err = np.linalg.norm(R - A) / np.linalg.norm(A)
    tol = 1e-10 if dt == "f64" else 1e-4
    print(f"hosvd {dt}: reconstruction rel err {err:.3e}", flush=True)
    assert err <= tol
    assert abs(float(torch.linalg.norm(core.double())) - np.linalg.norm(A)) <= tol * np.linalg.norm(A)


def test_hosvd_batch_uneven_layouts(gpu):
    import torch
    from ttm_amd.tucker import hosvd_batch, reconstruct_batch
    rng = np.random.default_rng(12)
    tensors, refs, ranks = [], [], []
    for t in range(6):
        n = tuple(int(x) for x in rng.integers(8, 48, size=3))
        r = tuple(int(min(x, nk)) for x, nk in zip(rng.integers(2, 12, size=3), n))
        pi = [int(x) + 1 for x in rng.permutation(3)]
        A = low_rank(rng, n, r)
        tensors.append(to_layout(torch, gpu, A, pi, torch.float64))
        refs.append(A)
        ranks.append(r)
    out = hosvd_batch(tensors, ranks)
    recs = reconstruct_batch([c for c, _ in out], [f for _, f in out])
    for R, A in zip(recs, refs):
        err = np.linalg.norm(R.cpu().numpy() - A) / np.linalg.norm(A)
        assert err <= 1e-10, err


@pytest.mark.parametrize("dtype", ["f32", "f64"])
def test_gram_matrices_are_this_librarys_ttm(gpu, dtype):
    """The HOSVD's Gram matrices A_(k) A_(k)^T come from this library's TTM
    (G = X x_2 X with X = A_(k)): equal to the float64 product of the
    unfolding with its transpose on the host, for every mode of tensors in
    several layouts, one with a long unfolding (the contraction split)."""
    import torch
    import ttm_amd
    from ttm_amd import tucker
    rng = np.random.default_rng(81)
    tdt = torch.float32 if dtype == "f32" else torch.float64
    for n, pi in (([40, 30, 20], [1, 2, 3]), ([17, 64, 9], [3, 1, 2]), ([256, 256, 64], [1, 2, 3])):
        X = rng.uniform(-1, 1, size=n)
        A = to_layout(torch, gpu, X, pi, tdt)
        Xq = A.double().cpu().numpy()   # the values the GPU sees (float32-rounded for f32)
        for k in range(3):
            G = tucker._gram(A, k)
            torch.cuda.synchronize()
            sym, splits = ttm_amd.last_kernel()
            U = np.moveaxis(Xq, k, 0).reshape(n[k], -1)
            ref = U @ U.T
            err = float(np.linalg.norm(G.cpu().numpy() - ref) / np.linalg.norm(ref))
            print(f"{n} pi={pi} mode {k + 1}: {sym} x{splits}, relative error {err:.2e}", flush=True)
            assert err <= 1e-13, (n, pi, k, err)
