"""Reduced-precision variants of the float32 product on the GPU.

Each variant is checked twice:
* against the CPU emulation of its own arithmetic (tests/lowp_emul.py),
  which it must match up to FP32 accumulation: |dC| <= 4 n_q 2^-24 sum|b||a|
  elementwise over the rounded operands;
* against the FP64 oracle (brute force of the definition), reporting max and
  normwise relative errors (the error table on A 64x48x32), with
  BF16X3 and FP16X2 held to the 1e-6 normwise target.
"""
import sys

import numpy as np
import pytest

import lowp_emul as le
import refcases as rc

pytestmark = pytest.mark.gpu

PRECS = ["bf16x3", "bf16", "fp16", "fp8", "fp16x2"]


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def run(torch, dev, q, a, na, wa, pia, b, nb, pib, nc, wc, precision, kernel=None):
    import ttm_amd
    A = torch.from_numpy(a).to(dev)
    B = torch.from_numpy(b).to(dev)
    C = torch.full((int(np.prod(nc)),), float("nan"), dtype=torch.float32, device=dev)
    ttm_amd.force_kernel(kernel)
    try:
        ttm_amd.ttm_raw(q, len(na), A, na, wa, pia, B, nb, pib, C, nc, wc, precision=precision)
        torch.cuda.synchronize()
    finally:
        ttm_amd.force_kernel(None)
    return C.cpu().numpy()


def folded(a, na, pia, q):
    """A as (outer, nq, w) in the folded order of its layout."""
    qh = pia.index(q) + 1
    w = int(np.prod([na[m - 1] for m in pia[: qh - 1]])) if qh > 1 else 1
    outer = int(np.prod([na[m - 1] for m in pia[qh:]])) if qh < len(pia) else 1
    return a.reshape(outer, na[q - 1], w)


def case(rng, na, pia, q, m, colmajor, dist):
    wa = rc.dense_strides(na, pia)
    n = int(np.prod(na))
    if dist == "uniform":
        a = rng.uniform(-1, 1, n).astype(np.float32)
        Bm = rng.uniform(-1, 1, (m, na[q - 1])).astype(np.float32)
    else:  # sign-random log-uniform magnitudes in [1e-3, 1e3]
        a = (rng.choice([-1, 1], n) * 10.0 ** rng.uniform(-3, 3, n)).astype(np.float32)
        Bm = (rng.choice([-1, 1], (m, na[q - 1])) * 10.0 ** rng.uniform(-3, 3, (m, na[q - 1]))).astype(np.float32)
    b = (Bm.T if colmajor else Bm).reshape(-1).copy()
    nc = list(na)
    nc[q - 1] = m
    wc = rc.dense_strides(nc, pia)
    return a, wa, Bm, b, [m, na[q - 1]], ([1, 2] if colmajor else [2, 1]), nc, wc


def check_vs_emulation(c, a, na, pia, q, Bm, prec, label, sliced=False):
    A3 = folded(a, na, pia, q)
    ref = le.emulate(prec, A3, Bm, sliced=sliced)        # (outer, m, w)
    mag = {"bf16": lambda x: np.abs(le.bf16(x)), "fp16": lambda x: np.abs(le.fp16(x)),
           "bf16x3": lambda x: np.abs(x.astype(np.float64)) * 1.001,
           "fp16x2": lambda x: np.abs(x.astype(np.float64)) * 1.001,
           "fp8": lambda x: np.abs(x.astype(np.float64)) * 1.2}[prec]
    # FP32 accumulation bound; the scaled FP8 instruction sums 128 products per
    # step inside the matrix core with undocumented intermediate rounding, so
    # it gets a 4x wider (still ~1000x tighter than FP8 rounding) allowance,
    # counted over the 128 terms of one instruction when n_q is shorter.
    terms = max(na[q - 1], 128) if prec == "fp8" else na[q - 1]
    factor = 16 if prec == "fp8" else 4
    bound = factor * terms * 2.0 ** -24 * np.einsum("jl,kli->kji", mag(Bm), mag(A3)) + 1e-30
    if prec == "fp8":
        # plus the smallest normal (8 subnormal steps) per operand and product:
        # the matrix core may flush e4m3fn subnormal inputs, which this
        # comparison does not pin (a wrong layout or scale errs by the values
        # themselves, orders of magnitude more)
        qa, qb = 8 * le.fp8_block_quantum(A3, axis=1), 8 * le.fp8_block_quantum(Bm, axis=1)
        bound = bound + np.einsum("jl,kli->kji", qb, mag(A3)) + np.einsum("jl,kli->kji", mag(Bm), qa)
    got = c.reshape(ref.shape)
    err = np.abs(got - ref)
    worst = float(np.max(err / bound))
    assert np.all(err <= bound), f"{label}: emulation mismatch, max err/bound {worst:.2f}"
    return worst


def errors_vs_fp64(c, c64):
    c = c.astype(np.float64)
    return (float(np.max(np.abs(c - c64)) / np.max(np.abs(c64))),
            float(np.linalg.norm(c - c64) / np.linalg.norm(c64)))


@pytest.mark.parametrize("prec", PRECS)
def test_variant_matches_its_emulation(gpu, prec):
    import torch
    import ttm_amd
    from ttm_amd._lib import PRECISION_NAMES
    rng = np.random.default_rng(11)
    shapes = [([64, 48, 32], [1, 2, 3], 2, 40), ([16, 24, 20], [3, 2, 1], 3, 64), ([40, 12, 8], [1, 2, 3], 1, 33),
              ([8, 36, 10], [2, 1, 3], 2, 130), ([20, 8, 5], [1, 3, 2], 2, 7), ([13, 7, 9], [1, 2, 3], 2, 5),
              ([256, 4, 6], [3, 1, 2], 1, 17),
              # more than 128 rows of B: several row blocks (ragged rows, columns and slices)
              ([12, 40, 24], [1, 2, 3], 2, 300), ([36, 20, 16], [1, 2, 3], 1, 260), ([64, 10, 14], [1, 2, 3], 1, 512),
              # NN view with n_q % 4 != 0: rows of a row-major B are not 16-byte aligned
              ([8, 7, 12], [1, 2, 3], 2, 20)]
    for na, pia, q, m in shapes:
        for colmajor in (True, False):
            for dist in ("uniform", "loguniform"):
                a, wa, Bm, b, nb, pib, nc, wc = case(rng, na, pia, q, m, colmajor, dist)
                for kernel in (None, "lowp_generic"):
                    c = run(torch, gpu, q, a, na, wa, pia, b, nb, pib, nc, wc, prec, kernel)
                    # the family that ran: FP16X2's tiled kernel scales A per block and
                    # slice in its 64-row blocks (at most 64 rows of B)
                    fam = kernel or ttm_amd.plan_kernel(q, na, pia, nb, pib, precision=int(PRECISION_NAMES[prec]))
                    # reorder C into the folded (outer, m, w) order of its layout
                    c3 = folded(c, nc, pia, q)
                    check_vs_emulation(c3, a, na, pia, q, Bm, prec, f"{prec} {na} {pia} q={q} m={m} {dist} {fam}",
                                       sliced=fam == "lowp" and m <= 64)


def test_variants_random_shapes_match_their_emulation(gpu):
    """Seeded random orders, extents (one of about 300 in half the cases:
    several 128-column blocks, partial slices), layouts, q, m (one or two row
    blocks, ragged) and B orientations for every reduced precision, each
    against the emulation of its own arithmetic (the bound of
    check_vs_emulation); plus FP8 on the NN view with n_q = 300 (partial
    128-wide scale slices) in both B layouts."""
    import torch
    import ttm_amd
    from ttm_amd._lib import PRECISION_NAMES
    rng = np.random.default_rng(31)
    cases = [([12, 300, 8], [1, 2, 3], 2, 70), ([12, 300, 8], [1, 2, 3], 2, 200)]
    while len(cases) < 26:
        p = int(rng.integers(2, 5))
        na = [int(rng.integers(2, 14)) for _ in range(p)]
        if rng.random() < 0.5:
            na[int(rng.integers(0, p))] = int(rng.integers(250, 330))
        if np.prod(na) > 400_000:
            continue
        pia = [int(x) + 1 for x in rng.permutation(p)]
        q = int(rng.integers(1, p + 1))
        m = int(rng.choice([1, 5, 17, 64, 65, 130, 200]))
        cases.append((na, pia, q, m))
    checked = 0
    for i, (na, pia, q, m) in enumerate(cases):
        for prec in PRECS:
            colmajor = bool((i + len(prec)) % 2)
            for cm in ((False, True) if i < 2 else (colmajor,)):
                a, wa, Bm, b, nb, pib, nc, wc = case(rng, na, pia, q, m, cm, "uniform")
                c = run(torch, gpu, q, a, na, wa, pia, b, nb, pib, nc, wc, prec)
                fam = ttm_amd.plan_kernel(q, na, pia, nb, pib, precision=int(PRECISION_NAMES[prec]))
                c3 = folded(c, nc, pia, q)
                label = f"{prec} {na} {pia} q={q} m={m} col={cm} {fam}"
                _say(label)
                check_vs_emulation(c3, a, na, pia, q, Bm, prec, label, sliced=fam == "lowp" and m <= 64)
                checked += 1
    assert checked >= 5 * 26


def test_precision_error_table_64x48x32(gpu, oracle):
    """A 64x48x32 first-order, q=2, B 40x48, float32 inputs, U[-1,1] and log-uniform."""
    import torch
    na, pia, q, m = [64, 48, 32], [1, 2, 3], 2, 40
    wa = rc.dense_strides(na, pia)
    nc = [64, 40, 32]
    wc = rc.dense_strides(nc, pia)
    table = {}
    for dist in ("uniform", "loguniform"):
        if dist == "uniform":
            a = oracle.fill_uniform(int(np.prod(na)), 0x77140003)
            b = oracle.fill_uniform(m * 48, 0x77140013)
        else:
            a = oracle.fill_loguniform(int(np.prod(na)), 0x77140023)
            b = oracle.fill_loguniform(m * 48, 0x77140033)
        c64 = oracle.ttm_brute_f64(q, a, na, wa, b, m, True, nc, wc)
        for prec in ["default"] + PRECS:
            c = run(torch, gpu, q, a, na, wa, pia, b, [m, 48], [1, 2], nc, wc, prec)
            mx, nw = errors_vs_fp64(c, c64)
            table[(dist, prec)] = (mx, nw)
            _say(f"64x48x32 {dist:10s} {prec:8s} max-rel {mx:.3e}  normwise-rel {nw:.3e}")
    for dist in ("uniform", "loguniform"):
        fp32 = table[(dist, "default")][1]
        assert table[(dist, "bf16x3")][1] <= 1e-6, table[(dist, "bf16x3")]
        assert table[(dist, "bf16x3")][1] <= 4 * fp32 + 1e-9
        assert table[(dist, "fp16x2")][1] <= 1e-6, table[(dist, "fp16x2")]
        assert 1e-5 < table[(dist, "bf16")][1] < 2e-2
        assert 1e-6 < table[(dist, "fp16")][1] < 3e-3
        assert 1e-3 < table[(dist, "fp8")][1] < 0.15


def test_fp16x2_infinity_sets_no_scale(gpu):
    """Data beyond binary16's range (|x| up to 1e6) with one infinity in A: the
    scale comes from the finite values, so every entry of C that does not
    involve the infinity matches the emulation, and those that do are not
    finite (a scale taken from the infinity would overflow every value)."""
    import torch
    rng = np.random.default_rng(21)
    na, pia, q, m = [64, 48, 32], [1, 2, 3], 2, 40
    for kernel in (None, "lowp_generic"):
        a, wa, Bm, b, nb, pib, nc, wc = case(rng, na, pia, q, m, False, "uniform")
        a = (a * 1.0e6).astype(np.float32)
        Bm = (Bm * 1.0e6).astype(np.float32)
        b = Bm.reshape(-1).copy()
        a[1234] = np.inf
        c = run(torch, gpu, q, a, na, wa, pia, b, nb, pib, nc, wc, "fp16x2", kernel)
        c3 = folded(c, nc, pia, q).astype(np.float64)
        A3 = folded(a, na, pia, q)
        with np.errstate(invalid="ignore"):
            ref = le.emulate("fp16x2", A3, Bm, sliced=kernel is None)  # m = 40: the 64-row blocks
        touched = np.einsum("jl,kli->kji", np.ones(Bm.shape), np.isinf(A3).astype(np.float64)) > 0
        assert touched.sum() == m and not np.isfinite(c3[touched]).any(), kernel
        A3f = np.where(np.isinf(A3), 0, A3).astype(np.float64)
        bound = 4 * na[q - 1] * 2.0 ** -24 * np.einsum("jl,kli->kji", np.abs(Bm.astype(np.float64)), np.abs(A3f)) * 1.001
        err = np.abs(c3 - ref)[~touched]
        assert np.all(err <= bound[~touched]), (kernel, float(np.max(err / bound[~touched])))


def test_fp8_non_finite_values_set_no_scale(gpu):
    """FP8 with one infinity and one NaN in A and one infinity in B: the slice
    scales come from the finite values (the emulation's rule), so every entry
    of C that involves none of them matches the emulation, and every entry
    that does is NaN (e4m3fn has no infinity).  A scale taken from an
    infinity would flush every finite value of its slice to zero."""
    import torch
    rng = np.random.default_rng(22)
    na, pia, q, m = [64, 300, 24], [1, 2, 3], 2, 40
    for layout, kernel in (([1, 2, 3], None), ([2, 1, 3], None), ([1, 2, 3], "lowp_generic")):
        pia = layout
        a, wa, Bm, b, nb, pib, nc, wc = case(rng, na, pia, q, m, False, "uniform")
        A3 = folded(a, na, pia, q)
        k0, l0, i0 = 3, 77, min(5, A3.shape[2] - 1)
        A3[k0, l0, i0] = np.inf
        A3[k0 + 1, 201, 0] = np.nan
        Bm[7, 130] = -np.inf
        b = Bm.reshape(-1).copy()  # A3 is a view of a: a holds the planted values too
        c = run(torch, gpu, q, a, na, wa, pia, b, nb, pib, nc, wc, "fp8", kernel)
        c3 = folded(c, nc, pia, q).astype(np.float64)
        with np.errstate(invalid="ignore"):
            ref = le.emulate("fp8", A3, Bm)
        bad_a = ~np.isfinite(A3)
        bad_b = ~np.isfinite(Bm)
        touched = (np.einsum("jl,kli->kji", np.ones(Bm.shape), bad_a.astype(np.float64)) > 0) | \
                  (np.einsum("jl,kli->kji", bad_b.astype(np.float64), np.ones(A3.shape)) > 0)
        assert np.isnan(c3[touched]).all(), (layout, kernel)
        assert np.isfinite(c3[~touched]).all(), (layout, kernel)
        qa = le.fp8_block_quantize(np.where(bad_a, 0, A3), axis=1)
        qb = le.fp8_block_quantize(np.where(bad_b, 0, Bm), axis=1)
        bound = 4 * na[q - 1] * 2.0 ** -24 * np.einsum("jl,kli->kji", np.abs(qb), np.abs(qa)) * 1.001
        err = np.abs(c3 - ref)[~touched]
        assert np.all(err <= bound[~touched] + 1e-30), (layout, kernel, float(np.max(err)))


@pytest.mark.parametrize("prec", ["bf16x3", "fp16x2"])
def test_split_variants_integer_data_is_exact(gpu, prec):
    """Integers below 2^24 split exactly into three bfloat16 parts, and the
    scaled integers of these cases into two binary16 parts: the closed form
    holds."""
    import torch
    for na, pia, q, colmajor in [([4, 8, 12], [1, 2, 3], 2, True), ([16, 12, 8], [3, 2, 1], 3, False),
                                 ([8, 8, 8], [2, 1, 3], 1, True)]:
        m = 2 * na[q - 1]
        a, wa = rc.ttm_init(q, na, pia)
        b, pib = rc.make_b(m, na[q - 1], "ones", colmajor, np.float32)
        nc = list(na)
        nc[q - 1] = m
        wc = rc.dense_strides(nc, pia)
        for kernel in (None, "lowp_generic"):
            c = run(torch, gpu, q, a.astype(np.float32), na, wa, pia, b, [m, na[q - 1]], pib, nc, wc, prec, kernel)
            assert np.array_equal(rc.as_nd(c, nc, wc), rc.closed_form(q, na, pia, m, "ones")), (na, pia, q, kernel)


@pytest.mark.parametrize("prec", ["bf16x3", "fp16x2"])
def test_split_variants_full_size_sampled(gpu, prec):
    """Compute-bound A 1024x1024x512 products (m = 1024) in BF16X3 and FP16X2:
    sampled entries against the float64 definition."""
    import torch
    import ttm_amd
    na = [1024, 1024, 512]
    A = torch.empty(int(np.prod(na)), dtype=torch.float32, device=gpu).uniform_(-1, 1)
    rng = np.random.default_rng(3)
    for pia, q in (([1, 2, 3], 2), ([3, 2, 1], 3), ([1, 2, 3], 3)):
        m = 1024
        Bt = torch.empty((m, na[q - 1]), dtype=torch.float32, device=gpu).uniform_(-1, 1)
        wa = rc.dense_strides(na, pia)
        nc = list(na)
        nc[q - 1] = m
        wc = rc.dense_strides(nc, pia)
        C = torch.empty(int(np.prod(nc)), dtype=torch.float32, device=gpu)
        ttm_amd.ttm_raw(q, 3, A, na, wa, pia, Bt, [m, na[q - 1]], [2, 1], C, nc, wc, precision=prec)
        torch.cuda.synchronize()
        p = 3
        idx = [rng.integers(0, nc[r], size=2048) for r in range(p)]
        offc = sum(idx[r] * wc[r] for r in range(p))
        base = sum(idx[r] * wa[r] for r in range(p) if r != q - 1)
        fib_idx = torch.from_numpy((base[:, None] + np.arange(na[q - 1])[None, :] * wa[q - 1]).astype(np.int64))
        fib = A[fib_idx.to(gpu)].double().cpu().numpy()
        brow = Bt.double().cpu().numpy()[idx[q - 1]]
        expect = np.sum(fib * brow, axis=1)
        got = C[torch.from_numpy(offc.astype(np.int64)).to(gpu)].double().cpu().numpy()
        rel = np.linalg.norm(got - expect) / np.linalg.norm(expect)
        _say(f"{prec} 1024x1024x512 pi={pia} q={q}: sampled normwise-rel {rel:.3e}")
        assert rel <= 1e-6
        del C
    torch.cuda.empty_cache()


@pytest.mark.parametrize("prec", ["bf16", "fp16", "fp8"])
def test_plain_variants_full_size_sampled(gpu, prec):
    """A 1024x1024x512 products (m = 1024; TN and NN views, row- and
    column-major B) in the plain reduced precisions: 1024 sampled entries of C
    against the emulation of the arithmetic on their fibres, to the bound of
    check_vs_emulation."""
    import torch
    import ttm_amd
    na = [1024, 1024, 512]
    A = torch.empty(int(np.prod(na)), dtype=torch.float32, device=gpu).uniform_(-1, 1)
    rng = np.random.default_rng(4)
    for pia, q, colmajor in (([1, 2, 3], 1, False), ([1, 2, 3], 2, False), ([3, 2, 1], 3, True), ([3, 2, 1], 1, True)):
        m, nq = 1024, na[q - 1]
        Bt = torch.empty((m, nq), dtype=torch.float32, device=gpu).uniform_(-1, 1)
        Bdev = Bt.t().contiguous().reshape(-1) if colmajor else Bt.reshape(-1)
        wa = rc.dense_strides(na, pia)
        nc = list(na)
        nc[q - 1] = m
        wc = rc.dense_strides(nc, pia)
        C = torch.empty(int(np.prod(nc)), dtype=torch.float32, device=gpu)
        ttm_amd.ttm_raw(q, 3, A, na, wa, pia, Bdev, [m, nq], [1, 2] if colmajor else [2, 1], C, nc, wc,
                        precision=prec)
        torch.cuda.synchronize()
        idx = [rng.integers(0, nc[r], size=1024) for r in range(3)]
        offc = sum(idx[r] * wc[r] for r in range(3))
        base = sum(idx[r] * wa[r] for r in range(3) if r != q - 1)
        fib_idx = torch.from_numpy((base[:, None] + np.arange(nq)[None, :] * wa[q - 1]).astype(np.int64))
        a_f = A[fib_idx.to(gpu)].cpu().numpy()
        b_r = Bt.cpu().numpy()[idx[q - 1]]
        if prec == "fp8":
            qa, qb = le.fp8_block_quantize(a_f, axis=1), le.fp8_block_quantize(b_r, axis=1)
        else:
            rnd = le.bf16 if prec == "bf16" else le.fp16
            qa, qb = rnd(a_f), rnd(b_r)
        expect = np.sum(qa * qb, axis=1)
        magp = np.abs(qa) * np.abs(qb)
        if prec == "fp8":
            magp = np.abs(a_f.astype(np.float64)) * 1.2 * np.abs(b_r.astype(np.float64)) * 1.2
        bound = (16 if prec == "fp8" else 4) * nq * 2.0 ** -24 * np.sum(magp, axis=1) + 1e-30
        if prec == "fp8":
            bound = bound + np.sum(8 * le.fp8_block_quantum(b_r, axis=1) * np.abs(a_f) * 1.2, axis=1) \
                + np.sum(np.abs(b_r) * 1.2 * 8 * le.fp8_block_quantum(a_f, axis=1), axis=1)
        got = C[torch.from_numpy(offc.astype(np.int64)).to(gpu)].double().cpu().numpy()
        err = np.abs(got - expect)
        _say(f"{prec} 1024x1024x512 pi={pia} q={q} colmajor={colmajor}: max err/bound {float(np.max(err / bound)):.3f}")
        assert np.all(err <= bound), (prec, pia, q, colmajor, float(np.max(err / bound)))
        del C
    torch.cuda.empty_cache()
