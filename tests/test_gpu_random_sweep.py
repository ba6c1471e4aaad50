"""Seeded random sweep: every kernel family, and the FP32-level split precisions
(FP16X2, BF16X3), on random orders, extents (one long mode in a third of the
cases), layouts, modes, m, B orientations and unaligned base pointers, against
the FP64 brute force of the oracle.

Bar: normwise ||C - C_64||_F / (||A||_F ||B||_F) <= (4 n_q + 8) eps of the
element type (the split precisions are held to float32's eps as well).
"""
import sys

import numpy as np
import pytest

import refcases as rc

pytestmark = pytest.mark.gpu

FAMILIES = [None, "generic", "nn_stream", "tn_stream", "gemm", "skinny"]


def test_random_sweep_all_families_vs_fp64(gpu, oracle):
    import torch
    import ttm_amd
    rng = np.random.default_rng(2024)
    checks, failures = 0, []
    for it in range(300):
        p = int(rng.integers(1, 5))
        na = [int(rng.integers(1, 40)) for _ in range(p)]
        if rng.random() < 0.3:
            na[int(rng.integers(0, p))] = int(rng.integers(100, 600))
        if np.prod(na) > 2_000_000:
            continue
        pia = [int(x) + 1 for x in rng.permutation(p)]
        q = int(rng.integers(1, p + 1))
        m = int(rng.choice([1, 3, 4, 7, 13, 16, 17, 31, 33, 40, 64, 65, 130]))
        colmajor = bool(rng.random() < 0.5)
        dtype = np.float64 if rng.random() < 0.35 else np.float32
        offset = int(rng.random() < 0.2)
        wa = rc.dense_strides(na, pia)
        a = rng.uniform(-1, 1, int(np.prod(na))).astype(dtype)
        Bm = rng.uniform(-1, 1, (m, na[q - 1])).astype(dtype)
        b = (Bm.T if colmajor else Bm).reshape(-1).copy()
        nb, pib = [m, na[q - 1]], ([1, 2] if colmajor else [2, 1])
        nc = list(na)
        nc[q - 1] = m
        wc = rc.dense_strides(nc, pia)
        c64 = oracle.ttm_brute_f64(q, a, na, wa, b, m, colmajor, nc, wc)
        tdt = torch.float32 if dtype == np.float32 else torch.float64

        def up(x):
            t = torch.zeros(x.size + offset, dtype=tdt, device=gpu)
            t[offset:] = torch.from_numpy(np.ascontiguousarray(x)).to(gpu)
            return t[offset:]

        A, B = up(a), up(b)
        precs = [None] + (["fp16x2", "bf16x3"] if dtype == np.float32 else [])
        for prec in precs:
            for fam in (FAMILIES if prec is None else [None, "lowp_generic"]):
                C = up(np.full(int(np.prod(nc)), np.nan, dtype=dtype))
                ttm_amd.force_kernel(fam)
                try:
                    ttm_amd.ttm_raw(q, p, A, na, wa, pia, B, nb, pib, C, nc, wc, precision=prec)
                    torch.cuda.synchronize()
                finally:
                    ttm_amd.force_kernel(None)
                err = rc.normwise_error(C.cpu().numpy(), c64, a, b)
                checks += 1
                if not err <= (4 * na[q - 1] + 8) * rc.eps(dtype):
                    failures.append((it, na, pia, q, m, colmajor, np.dtype(dtype).name, offset, fam, prec, err))
    print(f"random sweep: {checks} checks, {len(failures)} failures", flush=True)
    sys.stdout.flush()
    assert checks > 1500 and not failures, failures[:5]
