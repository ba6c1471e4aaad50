"""Under-filled grids and odd extents on the GPU (runs on an MI355X).

* GEMV-like products (q fastest with at most eight slabs, p = 1 included:
  the reference's case 1, include/tlib/detail/mtm.h:171,212) run the skinny
  kernels with the contraction split over the chip;
* long contractions with few blocks ([16, 16, 2^20] with q = 3) split the
  gemm tiles along the contraction;
* odd extents (w or n_q not a multiple of the vector width) and base
  pointers one element off 16 bytes run the tile kernels on element-aligned
  vectors instead of one thread per element.

Bars: small and medium shapes against the oracle's FP64 brute force,
normwise ||dC|| / (||A|| ||B||) <= 4 n_q eps; full-size shapes on sampled
entries against the float64 definition with the per-entry bound
|dC| <= 4 n_q eps ||a_fibre|| ||b_row|| and, because that bound is loose for
n_q of 2^20 and more, also |dC| <= 4 sqrt(n_q) eps ||a_fibre|| ||b_row||
(a dropped or doubled split range misses it by orders of magnitude).
"""
import sys

import numpy as np
import pytest

import refcases as rc

pytestmark = pytest.mark.gpu


def _say(*args):
    print(*args, flush=True)
    sys.stdout.flush()


def _tdt(torch, dtype):
    return torch.float32 if dtype == np.float32 else torch.float64


def _check_entries(got, expect, fib_norm, brow_norm, nq, dtype, label):
    eps = rc.eps(dtype)
    err = np.abs(got - expect)
    loose = 4 * nq * eps * fib_norm * brow_norm
    tight = 4 * np.sqrt(nq) * eps * fib_norm * brow_norm
    worst = float(np.max(err / np.maximum(tight, 1e-300)))
    _say(f"{label}: max err / (4 sqrt(n_q) eps |a||b|) = {worst:.3e}")
    assert np.all(err <= loose), label
    assert np.all(err <= tight), label


def _sampled(torch, A, na, wa, Bm, C, nc, wc, q, dtype, rng, label, samples=1024):
    """Sampled entries of C against the float64 definition (Bm is m x n_q)."""
    p = len(na)
    idx = [rng.integers(0, nc[r], size=samples) for r in range(p)]
    offc = sum(idx[r] * wc[r] for r in range(p))
    base_a = sum(idx[r] * wa[r] for r in range(p) if r != q - 1)
    nq = na[q - 1]
    got = C[torch.from_numpy(offc.astype(np.int64)).to(C.device)].double().cpu().numpy()
    expect = np.empty(samples)
    fn = np.empty(samples)
    bn = np.empty(samples)
    lq = torch.arange(nq, device=A.device, dtype=torch.int64)
    for s0 in range(0, samples, 64):  # fibres of A in batches (n_q may be 2^24)
        s1 = min(samples, s0 + 64)
        fidx = torch.from_numpy(base_a[s0:s1].astype(np.int64)).to(A.device)[:, None] + lq[None, :] * wa[q - 1]
        fib = A[fidx].double()
        brow = Bm[torch.from_numpy(idx[q - 1][s0:s1].astype(np.int64)).to(Bm.device)].double()
        expect[s0:s1] = (fib * brow).sum(dim=1).cpu().numpy()
        fn[s0:s1] = torch.linalg.norm(fib, dim=1).cpu().numpy()
        bn[s0:s1] = torch.linalg.norm(brow, dim=1).cpu().numpy()
    _check_entries(got, expect, fn, bn, nq, dtype, label)


def _run(torch, ttm_amd, q, A, na, pia, Bm, colmajor, dtype):
    nq = na[q - 1]
    m = Bm.shape[0]
    wa = rc.dense_strides(na, pia)
    nc = list(na)
    nc[q - 1] = m
    wc = rc.dense_strides(nc, pia)
    b = (Bm.t().contiguous() if colmajor else Bm.contiguous()).reshape(-1)
    C = torch.full((int(np.prod(nc)),), float("nan"), dtype=_tdt(torch, dtype), device=A.device)
    ttm_amd.ttm_raw(q, len(na), A, na, wa, pia, b, [m, nq], [1, 2] if colmajor else [2, 1], C, nc, wc)
    torch.cuda.synchronize()
    return C, wa, nc, wc


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_gemv_p1_vs_oracle(gpu, oracle, dtype):
    """p = 1, n_q = 2^24, m in {1, 16}, both B orientations: the skinny kernel
    with the contraction split, against the oracle's FP64 brute force."""
    import torch
    import ttm_amd
    nq = 1 << 24
    rng = np.random.default_rng(41)
    a = rng.uniform(-1, 1, nq).astype(dtype)
    A = torch.from_numpy(a).to(gpu)
    for m in (1, 16):
        for colmajor in (False, True):
            Bm_h = rng.uniform(-1, 1, (m, nq)).astype(dtype)
            Bm = torch.from_numpy(Bm_h).to(gpu)
            label = f"gemv {np.dtype(dtype).name} m={m} col={colmajor}"
            _say(label, ttm_amd.plan_kernel_symbol(1, [nq], [1], [m, nq], [1, 2] if colmajor else [2, 1],
                                                   is_double=dtype == np.float64))
            C, wa, nc, wc = _run(torch, ttm_amd, 1, A, [nq], [1], Bm, colmajor, dtype)
            sym, splits = ttm_amd.last_kernel()
            _say(f"{label}: ran {sym} with {splits} splits")
            assert sym.startswith("k_skinny_") and splits > 64, (sym, splits)
            c = C.cpu().numpy()
            b_flat = (Bm_h.T if colmajor else Bm_h).reshape(-1).copy()
            c64 = oracle.ttm_brute_f64(1, a, [nq], wa, b_flat, m, colmajor, nc, wc)
            assert rc.normwise_error(c, c64, a, b_flat) <= 4 * nq * rc.eps(dtype)
            fn = np.full(m, np.linalg.norm(a.astype(np.float64)))
            bn = np.linalg.norm(Bm_h.astype(np.float64), axis=1)
            _check_entries(c.astype(np.float64), c64, fn, bn, nq, dtype, label)
            del Bm, C
    torch.cuda.empty_cache()


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_skinny_2p20x4_q1_m1024(gpu, dtype):
    """A [2^20, 4] first-order, q = 1, m = 1024 (four slabs, B of 1024 x 2^20):
    the skinny kernels for both B orientations, sampled."""
    import torch
    import ttm_amd
    na, pia, q, m = [1 << 20, 4], [1, 2], 1, 1024
    rng = np.random.default_rng(42)
    A = torch.empty(int(np.prod(na)), dtype=_tdt(torch, dtype), device=gpu).uniform_(-1, 1)
    for colmajor in (True, False):
        Bm = torch.empty((m, na[0]), dtype=_tdt(torch, dtype), device=gpu).uniform_(-1, 1)
        label = f"skinny [2^20,4] {np.dtype(dtype).name} col={colmajor}"
        _say(label)
        C, wa, nc, wc = _run(torch, ttm_amd, q, A, na, pia, Bm, colmajor, dtype)
        sym, splits = ttm_amd.last_kernel()
        _say(f"{label}: ran {sym} with {splits} splits")
        assert sym.startswith("k_skinny_") and splits > 1, (sym, splits)
        _sampled(torch, A, na, wa, Bm, C, nc, wc, q, dtype, rng, label)
        del Bm, C
    torch.cuda.empty_cache()


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_split_gemm_16x16x2p20_q3(gpu, dtype):
    """A [16, 16, 2^20] first-order, q = 3 (256 columns, n_q = 2^20) at
    m = 16 and 1024: the gemm tiles split along the contraction, sampled."""
    import torch
    import ttm_amd
    na, pia, q = [16, 16, 1 << 20], [1, 2, 3], 3
    rng = np.random.default_rng(43)
    A = torch.empty(int(np.prod(na)), dtype=_tdt(torch, dtype), device=gpu).uniform_(-1, 1)
    for m, colmajor in ((16, False), (1024, True), (1024, False)):
        Bm = torch.empty((m, na[q - 1]), dtype=_tdt(torch, dtype), device=gpu).uniform_(-1, 1)
        label = f"split gemm [16,16,2^20] {np.dtype(dtype).name} m={m} col={colmajor}"
        _say(label)
        C, wa, nc, wc = _run(torch, ttm_amd, q, A, na, pia, Bm, colmajor, dtype)
        sym, splits = ttm_amd.last_kernel()
        _say(f"{label}: ran {sym} with {splits} splits")
        assert sym.startswith("k_gemm<") and splits >= 16, (sym, splits)
        _sampled(torch, A, na, wa, Bm, C, nc, wc, q, dtype, rng, label)
        del Bm, C
    torch.cuda.empty_cache()


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_odd_extents_full_size_sampled(gpu, dtype):
    """Extents that are not multiples of the vector width at full size: the
    padded NN tile (w = 999), TN tiles on rows of 1023 and 513 elements."""
    import torch
    import ttm_amd
    rng = np.random.default_rng(44)
    cases = [([999, 1000, 500], [1, 2, 3], 2, 16), ([999, 1000, 500], [1, 2, 3], 2, 1024),
             ([1023, 1024, 512], [1, 2, 3], 1, 16), ([1023, 1024, 512], [1, 2, 3], 1, 1024),
             ([1024, 1024, 513], [3, 2, 1], 3, 16)]
    for na, pia, q, m in cases:
        A = torch.empty(int(np.prod(na)), dtype=_tdt(torch, dtype), device=gpu).uniform_(-1, 1)
        for colmajor in (True, False):
            Bm = torch.empty((m, na[q - 1]), dtype=_tdt(torch, dtype), device=gpu).uniform_(-1, 1)
            label = f"odd {np.dtype(dtype).name} na={na} pi={pia} q={q} m={m} col={colmajor}"
            _say(label)
            C, wa, nc, wc = _run(torch, ttm_amd, q, A, na, pia, Bm, colmajor, dtype)
            sym, splits = ttm_amd.last_kernel()
            _say(f"{label}: ran {sym} x{splits}")
            assert not sym.startswith("k_generic"), sym
            _sampled(torch, A, na, wa, Bm, C, nc, wc, q, dtype, rng, label)
            del Bm, C
        del A
    torch.cuda.empty_cache()


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_offset_pointers_and_odd_extents_vs_oracle(gpu, oracle, dtype):
    """Base pointers one element past 16-byte alignment and odd extents, every
    view, both B orientations, against the FP64 brute force: no
    one-thread-per-element kernel runs."""
    import torch
    import ttm_amd
    rng = np.random.default_rng(45)
    cases = [([37, 41, 29], [1, 2, 3], 2, 19), ([37, 41, 29], [1, 2, 3], 1, 70), ([37, 41, 29], [3, 2, 1], 3, 5),
             ([3, 700, 9], [1, 2, 3], 2, 33), ([2, 5, 300], [1, 2, 3], 3, 130), ([64, 48, 32], [1, 2, 3], 2, 40),
             ([1500, 3], [1, 2], 1, 16), ([5000], [1], 1, 7), ([12, 7, 600], [3, 1, 2], 3, 64)]
    tdt = _tdt(torch, dtype)
    for na, pia, q, m in cases:
        for colmajor in (False, True):
            for offset in (0, 1):
                wa = rc.dense_strides(na, pia)
                a = rng.uniform(-1, 1, int(np.prod(na))).astype(dtype)
                Bm = rng.uniform(-1, 1, (m, na[q - 1])).astype(dtype)
                b = (Bm.T if colmajor else Bm).reshape(-1).copy()
                nc = list(na)
                nc[q - 1] = m
                wc = rc.dense_strides(nc, pia)

                def up(x):
                    t = torch.zeros(x.size + offset, dtype=tdt, device=gpu)
                    t[offset:] = torch.from_numpy(np.ascontiguousarray(x)).to(gpu)
                    return t[offset:]

                A, B = up(a), up(b)
                C = up(np.full(int(np.prod(nc)), np.nan, dtype=dtype))
                label = f"offset={offset} {np.dtype(dtype).name} na={na} pi={pia} q={q} m={m} col={colmajor}"
                _say(label)
                ttm_amd.ttm_raw(q, len(na), A, na, wa, pia, B, [m, na[q - 1]], [1, 2] if colmajor else [2, 1],
                                C, nc, wc)
                torch.cuda.synchronize()
                sym, splits = ttm_amd.last_kernel()
                assert not sym.startswith("k_generic"), (label, sym)
                c64 = oracle.ttm_brute_f64(q, a, na, wa, b, m, colmajor, nc, wc)
                err = rc.normwise_error(C.cpu().numpy(), c64, a, b)
                _say(f"{label}: {sym} x{splits} normwise {err:.3e}")
                assert err <= 4 * na[q - 1] * rc.eps(dtype), (label, sym, err)


def test_split_sweep_vs_oracle(gpu, oracle):
    """Seeded sweep of split-eligible shapes (q fastest with few slabs and a
    long contraction; NN views with few columns and a long contraction),
    every split count the tuner would try, against the FP64 brute force."""
    import torch
    import ttm_amd
    rng = np.random.default_rng(46)
    checks = 0
    for it in range(40):
        dtype = np.float64 if rng.random() < 0.4 else np.float32
        tdt = _tdt(torch, dtype)
        nq = int(rng.integers(512, 40000))
        if rng.random() < 0.5:   # TN: q fastest, outer <= 8
            outer = int(rng.integers(1, 9))
            na, pia, q = ([nq, outer] if outer > 1 else [nq]), ([1, 2] if outer > 1 else [1]), 1
        else:                    # NN: w columns below q, q slowest
            w = int(rng.integers(2, 70))
            na, pia, q = [w, nq], [1, 2], 2
        m = int(rng.choice([1, 2, 5, 16, 17, 40, 100]))
        colmajor = bool(rng.random() < 0.5)
        wa = rc.dense_strides(na, pia)
        a = rng.uniform(-1, 1, int(np.prod(na))).astype(dtype)
        Bm = rng.uniform(-1, 1, (m, nq)).astype(dtype)
        b = (Bm.T if colmajor else Bm).reshape(-1).copy()
        nc = list(na)
        nc[q - 1] = m
        wc = rc.dense_strides(nc, pia)
        c64 = oracle.ttm_brute_f64(q, a, na, wa, b, m, colmajor, nc, wc)
        A, B = torch.from_numpy(a).to(gpu), torch.from_numpy(b).to(gpu)
        C = torch.full((int(np.prod(nc)),), float("nan"), dtype=tdt, device=gpu)
        args = (q, len(na), A, na, wa, pia, B, [m, nq], [1, 2] if colmajor else [2, 1], C, nc, wc)
        label = f"sweep {it} {np.dtype(dtype).name} na={na} q={q} m={m} col={colmajor}"
        _say(label)
        splits = ttm_amd.tune_raw(*args)   # runs every candidate split count
        ttm_amd.ttm_raw(*args)             # the tuned choice
        torch.cuda.synchronize()
        sym, used = ttm_amd.last_kernel()
        assert used == splits or sym.startswith(("k_nn_stream", "k_tn_lds")), (label, sym, used, splits)
        err = rc.normwise_error(C.cpu().numpy(), c64, a, b)
        _say(f"{label}: {sym} x{used} normwise {err:.3e}")
        assert err <= 4 * nq * rc.eps(dtype), (label, sym, err)
        checks += 1
    ttm_amd.tune_clear()
    assert checks == 40


def test_tuned_split_is_reused_and_cleared(gpu):
    """ttm_stune stores its choice for the shape class: a later call of the
    same class runs it, tune_clear restores the planner's own choice."""
    import torch
    import ttm_amd
    na, pia, q, m = [1 << 22], [1], 1, 16
    A = torch.empty(na[0], device=gpu).uniform_(-1, 1)
    B = torch.empty((m, na[0]), device=gpu).uniform_(-1, 1)
    C = torch.empty(m, device=gpu)
    args = (q, 1, A, na, [1], pia, B, [m, na[0]], [2, 1], C, [m], [1])
    ttm_amd.tune_clear()
    ttm_amd.ttm_raw(*args)
    _, planned = ttm_amd.last_kernel()
    chosen = ttm_amd.tune_raw(*args)
    _say(f"planned {planned} splits, tuned {chosen}")
    ttm_amd.ttm_raw(*args)
    assert ttm_amd.last_kernel()[1] == chosen
    ttm_amd.tune_clear()
    ttm_amd.ttm_raw(*args)
    assert ttm_amd.last_kernel()[1] == planned
    torch.cuda.synchronize()


def test_host_entry_reuses_pooled_device_buffers(gpu):
    """Repeated synchronous host calls of one size take their device copies of
    A, B and C from the pool: it does not grow after the first call."""
    import ttm_amd
    na, pia, q, m = [96, 80, 64], [1, 2, 3], 2, 24
    rng = np.random.default_rng(47)
    a = rng.uniform(-1, 1, int(np.prod(na))).astype(np.float32)
    b = rng.uniform(-1, 1, m * na[q - 1]).astype(np.float32)
    nc = list(na)
    nc[q - 1] = m
    wa, wc = rc.dense_strides(na, pia), rc.dense_strides(nc, pia)
    c = np.empty(int(np.prod(nc)), dtype=np.float32)
    ttm_amd.ttm_raw(q, 3, a, na, wa, pia, b, [m, na[q - 1]], [2, 1], c, nc, wc)
    first = ttm_amd.workspace_bytes()
    for _ in range(10):
        ttm_amd.ttm_raw(q, 3, a, na, wa, pia, b, [m, na[q - 1]], [2, 1], c, nc, wc)
    _say(f"pool after the first call {first} B, after ten more {ttm_amd.workspace_bytes()} B")
    assert first > 0 and ttm_amd.workspace_bytes() == first


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_tn_tile_b_orientations_give_the_same_bits(gpu, oracle, dtype):
    """TN products (q fastest) on the 128-row tile: a row-major and a
    column-major B fill the tile's LDS with the same values (16-byte vectors
    along l or along j), so the two orientations give the same bits; each
    also against the FP64 brute force, with a split contraction (few column
    blocks, long n_q), ragged edges (n_q, m and the column count off the tile
    sizes) and an m that is not a multiple of the vector."""
    import torch
    import ttm_amd
    rng = np.random.default_rng(71)
    vec = 4 if dtype == np.float32 else 2
    cases = [([1000, 300], [1, 2], 1, 132),        # ragged n_q, columns and m
             ([8192, 40], [1, 2], 1, 256),         # 40 columns x 2 row blocks: contraction split
             ([96, 7, 130], [1, 3, 2], 1, 1024),   # three modes, q fastest
             ([200, 500], [1, 2], 1, 130 if vec == 4 else 131)]  # m % vec != 0
    for na, pia, q, m in cases:
        nq = na[q - 1]
        a = rng.uniform(-1, 1, int(np.prod(na))).astype(dtype)
        Bm = rng.uniform(-1, 1, (m, nq)).astype(dtype)
        A = torch.from_numpy(a).to(gpu)
        Bt = torch.from_numpy(Bm).to(gpu)
        C_row, wa, nc, wc = _run(torch, ttm_amd, q, A, na, pia, Bt, False, dtype)
        sym_row, splits = ttm_amd.last_kernel()
        C_col, _, _, _ = _run(torch, ttm_amd, q, A, na, pia, Bt, True, dtype)
        sym_col, _ = ttm_amd.last_kernel()
        label = f"{np.dtype(dtype).name} na={na} q={q} m={m}"
        _say(f"{label}: row-major B {sym_row} x{splits}, column-major B {sym_col}")
        assert sym_row.startswith("k_gemm<") and ", 4, 4, 2, 16, true," in sym_row, (label, sym_row)
        assert torch.equal(C_row, C_col), label
        c64 = oracle.ttm_brute_f64(q, a, na, wa, Bm.reshape(-1).copy(), m, False, nc, wc)
        err = rc.normwise_error(C_row.cpu().numpy(), c64, a, Bm.reshape(-1))
        _say(f"{label}: normwise {err:.3e}")
        assert err <= 4 * nq * rc.eps(dtype), (label, err)


def test_float64_tiles_of_17_to_64_rows_vs_oracle(gpu, oracle):
    """float64 products with 17..64 rows take the 128-column tiles (2 and 4
    row tiles of 16, three blocks per CU) in both views: against the FP64
    brute force, with ragged columns, slabs and contraction lengths."""
    import torch
    import ttm_amd
    rng = np.random.default_rng(91)
    cases = [([300, 70, 9], [1, 2, 3], 1, 17), ([300, 70, 9], [1, 2, 3], 2, 32), ([45, 33, 130], [3, 1, 2], 3, 33),
             ([45, 33, 130], [2, 3, 1], 1, 48), ([130, 257, 5], [1, 2, 3], 2, 64), ([64, 12, 200], [3, 2, 1], 3, 64),
             ([200, 9, 130], [1, 3, 2], 3, 40), ([45, 33, 13], [1, 2, 3], 3, 50), ([45, 33, 13], [1, 2, 3], 3, 20)]
    for na, pia, q, m in cases:   # the last two: odd slab width 1485 (the padded-slab tile)
        nq = na[q - 1]
        for colmajor in (False, True):
            a = rng.uniform(-1, 1, int(np.prod(na)))
            Bm = rng.uniform(-1, 1, (m, nq))
            A, Bt = torch.from_numpy(a).to(gpu), torch.from_numpy(Bm).to(gpu)
            C, wa, nc, wc = _run(torch, ttm_amd, q, A, na, pia, Bt, colmajor, np.float64)
            sym, splits = ttm_amd.last_kernel()
            label = f"na={na} pi={pia} q={q} m={m} col={colmajor}"
            _say(f"{label}: {sym} x{splits}")
            assert sym.startswith("k_gemm<double, ") and ", 2, 1, 16, " in sym, (label, sym)
            c64 = oracle.ttm_brute_f64(q, a, na, wa, Bm.reshape(-1).copy(), m, False, nc, wc)
            err = rc.normwise_error(C.cpu().numpy(), c64, a, Bm.reshape(-1))
            assert err <= 4 * nq * rc.eps(np.float64), (label, err)
