"""The allocation-free entry (ttm_sttm_ex_noalloc) against the entry it varies
(ttm_sttm_ex), through the C ABI on the GPU.

For each kind of scratch the library uses -- the partials of a contraction
split (gemm tiles and the skinny kernels), the converted B planes of BF16X3,
the scale exponents of FP16X2 and the quantized operands of FP8 -- two calls
in a row, the second on a smaller input, must give bit for bit the result of
ttm_sttm_ex on the same buffers (same kernels, same split counts, so the
same arithmetic in the same order).  A product whose scratch exceeds the
fixed array is rejected with TTM_ERR_ALLOC before anything runs.
"""
import ctypes
import sys

import numpy as np
import pytest

import refcases as rc

pytestmark = pytest.mark.gpu

SCRATCH_BYTES = 64 << 20


def _say(*args):
    print(*args, flush=True)
    sys.stdout.flush()


def _bind():
    import ttm_amd
    from ttm_amd._lib import _TTM_ARGS
    L = ttm_amd.lib()
    fn = L.ttm_sttm_ex_noalloc
    fn.restype = ctypes.c_int
    fn.argtypes = _TTM_ARGS + [ctypes.c_int, ctypes.c_void_p]
    return L, fn


def _call(fn, q, na, pia, A, B, nb, pib, C, prec, stream):
    from ttm_amd._lib import sizes
    nc = list(na)
    nc[q - 1] = nb[0]
    args = [ctypes.c_uint(q), ctypes.c_uint(len(na)), ctypes.c_void_p(A.data_ptr()), sizes(na),
            sizes(rc.dense_strides(na, pia)), sizes(pia), ctypes.c_void_p(B.data_ptr()), sizes(nb), sizes(pib),
            ctypes.c_void_p(C.data_ptr()), sizes(nc), sizes(rc.dense_strides(nc, pia)), ctypes.c_int(prec),
            ctypes.c_void_p(stream)]
    return fn(*args)


CASES = [
    # label, (na, pia, q, m, colmajor, precision) for the first and the smaller second call
    ("split gemm", ([16, 16, 1 << 18], [1, 2, 3], 3, 64, True, 0), ([16, 8, 1 << 16], [1, 2, 3], 3, 40, False, 0)),
    ("split skinny", ([1 << 22], [1], 1, 16, False, 0), ([1 << 20, 3], [1, 2], 1, 100, True, 0)),
    ("bf16x3 planes", ([128, 96, 64], [1, 2, 3], 2, 200, False, 1), ([64, 48, 32], [3, 1, 2], 1, 40, True, 1)),
    ("fp16x2 scales", ([256, 128, 64], [1, 2, 3], 2, 256, True, 5), ([40, 64, 36], [2, 3, 1], 3, 24, False, 5)),
    ("fp8 operands", ([256, 256, 64], [1, 2, 3], 1, 128, False, 4), ([96, 64, 32], [1, 2, 3], 2, 70, True, 4)),
]


@pytest.mark.parametrize("label,first,second", CASES, ids=[c[0] for c in CASES])
def test_noalloc_matches_pooled_entry(gpu, label, first, second):
    import torch
    import ttm_amd
    L, fn = _bind()
    stream = torch.cuda.current_stream().cuda_stream
    rng = np.random.default_rng(77)
    for call, (na, pia, q, m, colmajor, prec) in enumerate((first, second)):
        nq = na[q - 1]
        A = torch.from_numpy(rng.uniform(-1, 1, int(np.prod(na))).astype(np.float32)).to(gpu)
        Bm = torch.from_numpy(rng.uniform(-1, 1, (m, nq)).astype(np.float32)).to(gpu)
        B = (Bm.t().contiguous() if colmajor else Bm).reshape(-1)
        pib = [1, 2] if colmajor else [2, 1]
        nc_n = int(np.prod(na)) // nq * m
        C_ref = torch.full((nc_n,), float("nan"), device=gpu)
        C = torch.full((nc_n,), float("nan"), device=gpu)
        _say(f"{label} call {call + 1}: na={na} pi={pia} q={q} m={m} col={colmajor} precision={prec}")
        from ttm_amd._lib import check, sizes
        nc = list(na)
        nc[q - 1] = m
        check(L.ttm_sttm_ex(ctypes.c_uint(q), ctypes.c_uint(len(na)), ctypes.c_void_p(A.data_ptr()), sizes(na),
                            sizes(rc.dense_strides(na, pia)), sizes(pia), ctypes.c_void_p(B.data_ptr()),
                            sizes([m, nq]), sizes(pib), ctypes.c_void_p(C_ref.data_ptr()), sizes(nc),
                            sizes(rc.dense_strides(nc, pia)), ctypes.c_int(prec), ctypes.c_void_p(stream)))
        ref_kernel = ttm_amd.last_kernel()
        st = _call(fn, q, na, pia, A, B, [m, nq], pib, C, prec, stream)
        assert st == 0, (label, call, L.ttm_last_error().decode())
        torch.cuda.synchronize()
        assert ttm_amd.last_kernel() == ref_kernel, (ttm_amd.last_kernel(), ref_kernel)
        _say(f"{label} call {call + 1}: {ref_kernel[0]} x{ref_kernel[1]}")
        assert torch.equal(C, C_ref), (label, call)
        assert torch.isfinite(C).all()


def test_noalloc_rejects_scratch_beyond_the_array(gpu):
    """BF16X3 with B of 4096 x 4096: three bfloat16 planes (96 MiB) exceed the
    64 MiB array; the call fails with TTM_ERR_ALLOC and C is untouched."""
    import torch
    import ttm_amd
    L, fn = _bind()
    na, pia, q, m = [16, 4096], [1, 2], 2, 4096
    A = torch.empty(int(np.prod(na)), device=gpu).uniform_(-1, 1)
    B = torch.empty(m * na[q - 1], device=gpu).uniform_(-1, 1)
    C = torch.full((16 * m,), float("nan"), device=gpu)
    st = _call(fn, q, na, pia, A, B, [m, na[q - 1]], [2, 1], C, 1, torch.cuda.current_stream().cuda_stream)
    msg = L.ttm_last_error().decode()
    _say(f"rejected: status {st}: {msg}")
    assert st == int(ttm_amd.Status.ALLOC) and "scratch" in msg and str(SCRATCH_BYTES) in msg
    torch.cuda.synchronize()
    assert torch.isnan(C).all()
