"""Parity of the HIP path with the oracle (runs on an MI355X).

Every comparison goes through the C ABI (ttm_amd.ttm_raw -> libttm_amd.so).
Bars:
* closed-form sweeps (integer-valued data): bit-exact;
* random data: normwise ||C - C_ref||_F / (||A||_F ||B||_F) <= 4 n_q eps
  against the oracle's restatement of the reference (same precision) and
  against the FP64 brute force;
* full-size shapes: sampled entries against the float64 definition with the
  per-entry dot-product bound |dC| <= 4 n_q eps ||a_fibre|| ||b_row||.
"""
import json
import pathlib
import subprocess
import sys

import numpy as np
import pytest

import refcases as rc

pytestmark = pytest.mark.gpu

ROOT = pathlib.Path(__file__).resolve().parents[1]
GOLD = json.loads((ROOT / "tests" / "golden" / "reference_known_answers.json").read_text())
FAMILIES = ["generic", "nn_stream", "tn_stream", "gemm", "skinny"]


def _say(*args):
    print(*args, flush=True)
    sys.stdout.flush()


def run_gpu(torch, dev, q, a, na, wa, pia, b, nb, pib, nc, wc, kernel=None, offset=0):
    """C = A x_q B through the C ABI on device buffers; `offset` elements of
    padding in front of every buffer make the base pointers unaligned."""
    import ttm_amd

    def up(x):
        t = torch.zeros(x.size + offset, dtype=torch.float32 if x.dtype == np.float32 else torch.float64,
                        device=dev)
        t[offset:] = torch.from_numpy(np.ascontiguousarray(x)).to(dev)
        return t[offset:]

    A = up(a)
    B = up(b)
    C = up(np.full(int(np.prod(nc)), np.nan, dtype=a.dtype))
    ttm_amd.force_kernel(kernel)
    try:
        ttm_amd.ttm_raw(q, len(na), A, na, wa, pia, B, nb, pib, C, nc, wc)
        torch.cuda.synchronize()
    finally:
        ttm_amd.force_kernel(None)
    return C.cpu().numpy()


def random_case(rng, na, pia, q, m, colmajor, dtype):
    wa = rc.dense_strides(na, pia)
    a = rng.uniform(-1, 1, size=int(np.prod(na))).astype(dtype)
    Bm = rng.uniform(-1, 1, size=(m, na[q - 1])).astype(dtype)
    b = (Bm.T if colmajor else Bm).reshape(-1).copy()
    nc = list(na)
    nc[q - 1] = m
    wc = rc.dense_strides(nc, pia)
    return a, wa, b, [m, na[q - 1]], ([1, 2] if colmajor else [2, 1]), nc, wc


# ----------------------------------------------------------------------------- closed forms
def test_closed_form_sweep_exact(gpu, oracle):
    import torch
    n = 0
    for dtype, kinds in ((np.float64, ("ones", "jl")), (np.float32, ("ones",))):
        for na, pia, q, colmajor, bk in rc.closed_form_cases(orders=(2, 3, 4), b_kinds=kinds):
            m = 2 * na[q - 1]
            a, wa = rc.ttm_init(q, na, pia)
            b, pib = rc.make_b(m, na[q - 1], bk, colmajor, dtype)
            nc = list(na)
            nc[q - 1] = m
            wc = rc.dense_strides(nc, pia)
            c = run_gpu(torch, gpu, q, a.astype(dtype), na, wa, pia, b, [m, na[q - 1]], pib, nc, wc)
            expect = rc.closed_form(q, na, pia, m, bk)
            assert np.array_equal(rc.as_nd(c, nc, wc), expect), (na, pia, q, colmajor, bk, dtype)
            n += 1
    _say("closed-form cases", n)
    assert n > 10000


@pytest.mark.parametrize("family", FAMILIES)
def test_closed_form_each_family(gpu, family):
    """Every kernel family on the shapes it accepts (others fall back, still checked)."""
    import torch
    cases = list(rc.closed_form_cases(orders=(2, 3, 4), start=2, steps=3, b_kinds=("jl",)))[::5]
    cases += [([4, 8, 12], [1, 2, 3], 2, True, "jl"), ([8, 4, 16], [2, 1, 3], 2, False, "jl"),
              ([16, 12, 8], [3, 2, 1], 3, True, "jl"), ([16, 12, 8], [3, 2, 1], 3, False, "jl")]
    for na, pia, q, colmajor, bk in cases:
        m = 2 * na[q - 1] + 3
        a, wa = rc.ttm_init(q, na, pia)
        b, pib = rc.make_b(m, na[q - 1], bk, colmajor)
        nc = list(na)
        nc[q - 1] = m
        wc = rc.dense_strides(nc, pia)
        c = run_gpu(torch, gpu, q, a, na, wa, pia, b, [m, na[q - 1]], pib, nc, wc, kernel=family)
        assert np.array_equal(rc.as_nd(c, nc, wc), rc.closed_form(q, na, pia, m, bk)), (family, na, pia, q)


# ----------------------------------------------------------------------------- random data
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_64x48x32_q2_vs_oracle(gpu, oracle, dtype):
    """A 64x48x32 first-order, q = 2, B 40x48 (column- and row-major), uniform[-1,1]."""
    import torch
    for colmajor in (True, False):
        a = oracle.fill_uniform(64 * 48 * 32, 0x77140000 + 1, dtype)
        bm = oracle.fill_uniform(40 * 48, 0x77140000 + 2, dtype).reshape(40, 48)
        b = (bm.T if colmajor else bm).reshape(-1).copy()
        na, pia, q = [64, 48, 32], [1, 2, 3], 2
        wa = rc.dense_strides(na, pia)
        nc = [64, 40, 32]
        wc = rc.dense_strides(nc, pia)
        pib = [1, 2] if colmajor else [2, 1]
        c_ref, branch = oracle.ttm_reference(q, a, na, wa, pia, b, [40, 48], pib, nc, wc)
        c64 = oracle.ttm_brute_f64(q, a, na, wa, b, 40, colmajor, nc, wc)
        for fam in [None] + FAMILIES:
            c = run_gpu(torch, gpu, q, a, na, wa, pia, b, [40, 48], pib, nc, wc, kernel=fam)
            bound = 4 * 48 * rc.eps(dtype)
            e_ref = rc.normwise_error(c, c_ref, a, b)
            e_64 = rc.normwise_error(c, c64, a, b)
            _say(f"64x48x32 {np.dtype(dtype).name} colmajor={colmajor} kernel={fam}: vs ref {e_ref:.2e} "
                 f"vs fp64 {e_64:.2e} bound {bound:.2e}")
            assert e_ref <= bound and e_64 <= bound


SHAPES = [
    # (na, pia, q, m): odd extents, w not a multiple of 4, n_q odd, m not a multiple of 16 ...
    ([5, 7, 3], [1, 2, 3], 2, 9),
    ([5, 7, 3], [3, 1, 2], 1, 17),
    ([64, 33, 8], [1, 2, 3], 2, 64),
    ([12, 20, 8], [2, 1, 3], 3, 70),
    ([16, 16, 16], [3, 2, 1], 3, 5),
    ([16, 16, 16], [3, 2, 1], 1, 130),
    ([8, 12, 4, 6], [4, 2, 1, 3], 2, 33),
    ([6, 4, 5, 4, 3], [5, 1, 3, 2, 4], 4, 64),
    ([37], [1], 1, 11),
    ([20, 24], [1, 2], 1, 8), ([20, 24], [1, 2], 2, 8), ([20, 24], [2, 1], 1, 8), ([20, 24], [2, 1], 2, 8),
    ([1, 9, 1], [1, 2, 3], 2, 3),
    ([4, 1, 6], [2, 3, 1], 2, 7),
    ([300, 1, 5], [1, 2, 3], 1, 1),
]


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_random_shapes_all_families_vs_fp64(gpu, oracle, dtype):
    import torch
    rng = np.random.default_rng(1234)
    for na, pia, q, m in SHAPES:
        for colmajor in (True, False):
            a, wa, b, nb, pib, nc, wc = random_case(rng, na, pia, q, m, colmajor, dtype)
            c64 = oracle.ttm_brute_f64(q, a, na, wa, b, m, colmajor, nc, wc)
            bound = 4 * na[q - 1] * rc.eps(dtype)
            for fam in [None] + FAMILIES:
                for offset in (0, 1):
                    c = run_gpu(torch, gpu, q, a, na, wa, pia, b, nb, pib, nc, wc, kernel=fam, offset=offset)
                    e = rc.normwise_error(c, c64, a, b)
                    assert e <= bound, (na, pia, q, m, colmajor, fam, offset, e)


@pytest.mark.parametrize("layout", ["seeded", "bench", "first"])
def test_order5_random_layout_all_modes(gpu, oracle, layout):
    """A 32x24x40x16x48 float32, q = 1..5, B 64 x n_q, in a seeded random
    layout, in the layout the benchmark times (4, 2, 3, 1, 5) and in the
    first-order layout: sampled entries against the float64 definition (the
    full oracle is too slow at this size)."""
    import torch
    import ttm_amd
    na = [32, 24, 40, 16, 48]
    rng = np.random.default_rng(0x77140003)
    pia = {"seeded": [int(x) + 1 for x in rng.permutation(5)], "bench": [4, 2, 3, 1, 5],
           "first": [1, 2, 3, 4, 5]}[layout]
    A = torch.empty(int(np.prod(na)), dtype=torch.float32, device=gpu).uniform_(-1, 1)
    wa = rc.dense_strides(na, pia)
    for q in range(1, 6):
        m = 64
        Bt = torch.empty((m, na[q - 1]), dtype=torch.float32, device=gpu).uniform_(-1, 1)
        nc = list(na)
        nc[q - 1] = m
        wc = rc.dense_strides(nc, pia)
        C = torch.empty(int(np.prod(nc)), dtype=torch.float32, device=gpu)
        ttm_amd.ttm_raw(q, 5, A, na, wa, pia, Bt, [m, na[q - 1]], [2, 1], C, nc, wc)
        torch.cuda.synchronize()
        _check_sampled(torch, A, na, wa, Bt, C, nc, wc, q, np.float32, rng, f"order5 pi={pia} q={q}")
        # the same product in BF16X3: normwise-relative error of sampled entries <= 1e-6
        ttm_amd.ttm_raw(q, 5, A, na, wa, pia, Bt, [m, na[q - 1]], [2, 1], C, nc, wc, precision="bf16x3")
        torch.cuda.synchronize()
        rel = _sampled_normwise(torch, A, na, wa, Bt, C, nc, wc, q, rng)
        _say(f"order5 pi={pia} q={q} bf16x3: sampled normwise-rel {rel:.3e}")
        assert rel <= 1e-6, (pia, q, rel)


def _sampled_normwise(torch, A, na, wa, Bt, C, nc, wc, q, rng, samples=4096):
    """Normwise relative error of sampled entries of C against the float64 definition."""
    p = len(na)
    idx = [rng.integers(0, nc[r], size=samples) for r in range(p)]
    offc = sum(idx[r] * wc[r] for r in range(p))
    base_a = sum(idx[r] * wa[r] for r in range(p) if r != q - 1)
    lq = np.arange(na[q - 1])
    fib_idx = torch.from_numpy((base_a[:, None] + lq[None, :] * wa[q - 1]).astype(np.int64)).to(A.device)
    fib = A[fib_idx].double().cpu().numpy()
    brow = Bt.double().cpu().numpy()[idx[q - 1]]
    expect = np.sum(fib * brow, axis=1)
    got = C[torch.from_numpy(offc.astype(np.int64)).to(C.device)].double().cpu().numpy()
    return float(np.linalg.norm(got - expect) / np.linalg.norm(expect))


def _check_sampled(torch, A, na, wa, Bt, C, nc, wc, q, dtype, rng, label, samples=4096):
    """Sampled entries of C against the float64 definition with the dot-product bound."""
    p = len(na)
    idx = [rng.integers(0, nc[r], size=samples) for r in range(p)]
    offc = sum(idx[r] * wc[r] for r in range(p))
    base_a = sum(idx[r] * wa[r] for r in range(p) if r != q - 1)
    lq = np.arange(na[q - 1])
    fib_idx = torch.from_numpy((base_a[:, None] + lq[None, :] * wa[q - 1]).astype(np.int64)).to(A.device)
    fib = A[fib_idx].double().cpu().numpy()                      # samples x n_q
    brow = Bt.double().cpu().numpy()[idx[q - 1]]                 # samples x n_q
    expect = np.sum(fib * brow, axis=1)
    got = C[torch.from_numpy(offc.astype(np.int64)).to(C.device)].double().cpu().numpy()
    bound = 4 * na[q - 1] * rc.eps(dtype) * np.linalg.norm(fib, axis=1) * np.linalg.norm(brow, axis=1)
    err = np.abs(got - expect)
    worst = float(np.max(err / np.maximum(bound, 1e-300)))
    _say(f"{label}: sampled max err/bound = {worst:.3f}")
    assert np.all(err <= bound), label


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_1024x1024x512_sweep_sampled(gpu, dtype):
    """A 1024x1024x512, first- and last-order, q = 1..3, m in {16, 1024} (the bench workload)."""
    import torch
    import ttm_amd
    na = [1024, 1024, 512]
    tdt = torch.float32 if dtype == np.float32 else torch.float64
    rng = np.random.default_rng(99)
    A = torch.empty(int(np.prod(na)), dtype=tdt, device=gpu).uniform_(-1, 1)
    for pia in ([1, 2, 3], [3, 2, 1]):
        wa = rc.dense_strides(na, pia)
        for q in (1, 2, 3):
            for m in (16, 1024):
                Bt = torch.empty((m, na[q - 1]), dtype=tdt, device=gpu).uniform_(-1, 1)
                nc = list(na)
                nc[q - 1] = m
                wc = rc.dense_strides(nc, pia)
                C = torch.empty(int(np.prod(nc)), dtype=tdt, device=gpu)
                _say(f"1024x1024x512 {np.dtype(dtype).name} pi={pia} q={q} m={m} kernel="
                     f"{ttm_amd.plan_kernel(q, na, pia, [m, na[q - 1]], [2, 1], is_double=dtype == np.float64)}")
                ttm_amd.ttm_raw(q, 3, A, na, wa, pia, Bt, [m, na[q - 1]], [2, 1], C, nc, wc)
                torch.cuda.synchronize()
                _check_sampled(torch, A, na, wa, Bt, C, nc, wc, q, dtype, rng,
                               f"1024x1024x512 {np.dtype(dtype).name} pi={pia} q={q} m={m}", samples=1024)
                del C, Bt
    del A
    torch.cuda.empty_cache()


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_1024x1024x512_stream_kernels_sampled(gpu, dtype):
    """The streaming kernels forced on the bench extents (A 1024x1024x512, first- and
    last-order, q = 1..3): m = 16 with a row-major B and m = 40 with a column-major B."""
    import torch
    import ttm_amd
    na = [1024, 1024, 512]
    tdt = torch.float32 if dtype == np.float32 else torch.float64
    rng = np.random.default_rng(98)
    A = torch.empty(int(np.prod(na)), dtype=tdt, device=gpu).uniform_(-1, 1)
    for pia in ([1, 2, 3], [3, 2, 1]):
        wa = rc.dense_strides(na, pia)
        for q in (1, 2, 3):
            w = int(np.prod([na[r - 1] for r in pia[:pia.index(q)]]))
            fam = "tn_stream" if w == 1 else "nn_stream"
            for m, colmajor in ((16, False), (40, True)):
                Bm = torch.empty((m, na[q - 1]), dtype=tdt, device=gpu).uniform_(-1, 1)
                b = Bm.t().contiguous() if colmajor else Bm
                pib = [1, 2] if colmajor else [2, 1]
                nc = list(na)
                nc[q - 1] = m
                wc = rc.dense_strides(nc, pia)
                C = torch.empty(int(np.prod(nc)), dtype=tdt, device=gpu)
                ttm_amd.force_kernel(fam)
                try:
                    used = ttm_amd.plan_kernel(q, na, pia, [m, na[q - 1]], pib, is_double=dtype == np.float64)
                    _say(f"stream {np.dtype(dtype).name} pi={pia} q={q} m={m} col={colmajor} kernel={used}")
                    assert used == fam
                    ttm_amd.ttm_raw(q, 3, A, na, wa, pia, b, [m, na[q - 1]], pib, C, nc, wc)
                    torch.cuda.synchronize()
                finally:
                    ttm_amd.force_kernel(None)
                _check_sampled(torch, A, na, wa, Bm, C, nc, wc, q, dtype, rng,
                               f"stream {np.dtype(dtype).name} pi={pia} q={q} m={m}", samples=1024)
                del C, Bm, b
    del A
    torch.cuda.empty_cache()


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_tn_lds_edges_vs_oracle(gpu, oracle, dtype):
    """The TN streaming kernel with all of B in LDS (q fastest, m <= 16): ragged
    512-byte row segments, row counts that are not a multiple of 16, fewer
    tiles than waves, an odd number of segments per wave, m from 1 to 16, both
    B layouts and every B capacity (129, 257 and, for float64, 513 pieces per
    row), against the float64 brute force."""
    import torch
    import ttm_amd
    rng = np.random.default_rng(77)
    cases = [([4, 3, 5], 4), ([132, 7, 3], 16), ([516, 5, 3], 13), ([1024, 33, 2], 16), ([1020, 17, 1], 1),
             ([8, 1, 1], 3), ([260, 70, 9], 16), ([384, 100, 4], 7), ([132, 300, 70], 16), ([128, 512, 64], 12)]
    q, pia = 1, [1, 2, 3]  # q is the fastest mode: the TN view
    for na, m in cases:
        for colmajor in (False, True):
            a, wa, b, nb, pib, nc, wc = random_case(rng, na, pia, q, m, colmajor, dtype)
            ttm_amd.force_kernel("tn_stream")
            try:
                sym = ttm_amd.plan_kernel_symbol(q, na, pia, nb, pib, is_double=dtype == np.float64)
            finally:
                ttm_amd.force_kernel(None)
            assert sym.startswith("k_tn_lds<" + ("double" if dtype == np.float64 else "float")), sym
            c = run_gpu(torch, gpu, q, a, na, wa, pia, b, nb, pib, nc, wc, kernel="tn_stream")
            c64 = oracle.ttm_brute_f64(q, a, na, wa, b, m, colmajor, nc, wc)
            err = rc.normwise_error(c, c64, a, b)
            _say(f"tn_lds na={na} m={m} col={colmajor} {sym}: err {err:.2e}")
            assert err <= 4 * na[q - 1] * rc.eps(dtype), (na, m, colmajor)


# ----------------------------------------------------------------------------- python bindings
def test_ttmpy_reference_vectors_numpy_and_torch(gpu):
    """ttmpy/tests/test.py: exact equality with numpy.einsum, for ttm and all three ttms orders."""
    import torch
    from ttm_amd import ttmpy as tp
    for case in GOLD["ttmpy_einsum"]["ttm"]:
        A = np.arange(np.prod(case["a_shape"]), dtype=np.float64).reshape(case["a_shape"])
        B = np.arange(np.prod(case["b_shape"]), dtype=np.float64).reshape(case["b_shape"])
        D = np.einsum(case["einsum"], A, B)
        assert np.all(tp.ttm(case["q"], A, B) == D)
        Ct = tp.ttm(case["q"], torch.from_numpy(A).to(gpu), torch.from_numpy(B).to(gpu))
        assert Ct.is_cuda and np.all(Ct.cpu().numpy() == D)
        C32 = tp.ttm(case["q"], A.astype(np.float32), B.astype(np.float32))
        assert C32.dtype == np.float32 and np.all(C32 == D.astype(np.float32))
    for case in GOLD["ttmpy_einsum"]["ttms"]:
        A = np.arange(np.prod(case["a_shape"]), dtype=np.float64).reshape(case["a_shape"])
        Bs = [np.arange(np.prod(s), dtype=np.float64).reshape(s) for s in case["b_shapes"]]
        D = A
        for spec, Bm in zip(case["einsums"], Bs):
            D = np.einsum(spec, D, Bm)
        for order in GOLD["ttmpy_einsum"]["orders"]:
            assert np.all(tp.ttms(case["q"], A, Bs, order) == D), (case["q"], order)
            Ct = tp.ttms(case["q"], torch.from_numpy(A).to(gpu), [torch.from_numpy(x).to(gpu) for x in Bs], order)
            assert np.all(Ct.cpu().numpy() == D)


def test_ttmpy_errors_mirror_reference(gpu):
    from ttm_amd import ttmpy as tp
    with pytest.raises(ValueError, match="contraction mode should be greater than zero"):
        tp.ttm(4, np.zeros((2, 3, 4)), np.zeros((2, 2)))
    with pytest.raises(ValueError, match="second input should be a mtrix"):
        tp.ttm(1, np.zeros((2, 3, 4)), np.zeros(2))
    with pytest.raises(RuntimeError, match="contraction dimension of A and B are not equal"):
        tp.ttm(1, np.zeros((2, 3, 4)), np.zeros((5, 3)))
    with pytest.raises(RuntimeError, match="stride vector of A is not valid"):
        tp.ttm(1, np.asfortranarray(np.zeros((2, 3, 4))), np.zeros((5, 2)))
    with pytest.raises(ValueError, match="multiplication order"):
        tp.ttms(1, np.zeros((2, 3, 4)), [np.zeros((2, 3)), np.zeros((2, 4))], "sideways")
    with pytest.raises(ValueError, match="number of input matrices"):
        tp.ttms(1, np.zeros((2, 3, 4)), [np.zeros((2, 3))])


def test_any_dense_layout_tensor_api(gpu):
    """ttm_amd.ttm infers the layout from strides (permuted torch views, Fortran arrays)."""
    import torch
    import ttm_amd
    rng = np.random.default_rng(5)
    X = rng.uniform(-1, 1, size=(6, 10, 8))
    Bm = rng.uniform(-1, 1, size=(12, 10))
    D = np.einsum("ijk,xj->ixk", X, Bm)
    for arr in (X, np.asfortranarray(X)):
        C = ttm_amd.ttm(2, arr, Bm)
        np.testing.assert_allclose(C, D, rtol=1e-12, atol=1e-12)
    Xt = torch.from_numpy(np.ascontiguousarray(X.transpose(2, 0, 1))).to(gpu).permute(1, 2, 0)  # dense, permuted
    for b_layout in ("row", "col"):
        C = ttm_amd.ttm(2, Xt, torch.from_numpy(Bm).to(gpu), b_layout=b_layout)
        assert C.stride() == (12, 1, 72)  # C keeps A's layout pi = (2, 1, 3)
        np.testing.assert_allclose(C.cpu().numpy(), D, rtol=1e-12, atol=1e-12)


def test_host_entry_points_with_host_buffers(gpu, oracle):
    """The synchronous host entries (the drop-in semantics) on NumPy buffers."""
    import ttm_amd
    rng = np.random.default_rng(8)
    for dtype in (np.float32, np.float64):
        na, pia, q, m = [9, 14, 6], [2, 3, 1], 3, 21
        a, wa, b, nb, pib, nc, wc = random_case(rng, na, pia, q, m, True, dtype)
        c = np.full(int(np.prod(nc)), np.nan, dtype=dtype)
        ttm_amd.ttm_raw(q, 3, a, na, wa, pia, b, nb, pib, c, nc, wc)
        c64 = oracle.ttm_brute_f64(q, a, na, wa, b, m, True, nc, wc)
        assert rc.normwise_error(c, c64, a, b) <= 4 * na[q - 1] * rc.eps(dtype)


def test_host_entry_staging_of_large_buffers(gpu):
    """Host buffers of 16 MiB and more move through the pinned chunk pipeline
    (64 MiB chunks, ragged last chunk) both ways: the host entry must give the
    device entry's result bit for bit.  A is 150 MiB (three chunks, the last
    22 MiB), C is 176 MiB, B stays below the staging threshold."""
    import torch
    import ttm_amd
    na, pia, q, m = [256, 384, 400], [1, 2, 3], 1, 300
    rng = np.random.default_rng(12)
    a = rng.uniform(-1, 1, size=int(np.prod(na))).astype(np.float32)
    b = rng.uniform(-1, 1, size=m * na[q - 1]).astype(np.float32)
    nc = list(na)
    nc[q - 1] = m
    wa, wc = ttm_amd.dense_strides(na, pia), ttm_amd.dense_strides(nc, pia)
    c = np.full(int(np.prod(nc)), np.nan, dtype=np.float32)
    ttm_amd.ttm_raw(q, 3, a, na, wa, pia, b, [m, na[q - 1]], [1, 2], c, nc, wc)
    ad, bd = torch.from_numpy(a).to(gpu), torch.from_numpy(b).to(gpu)
    cd = torch.empty(c.size, dtype=torch.float32, device=gpu)
    ttm_amd.ttm_raw(q, 3, ad, na, wa, pia, bd, [m, na[q - 1]], [1, 2], cd, nc, wc)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(c, cd.cpu().numpy())
    # A's chunks land whole: one column of the product checked in float64
    i3 = 399
    ref = (b.reshape(na[0], m).T.astype(np.float64) @
           a.reshape(na[2], na[1], na[0])[i3].T.astype(np.float64))
    got = c.reshape(nc[2], nc[1], nc[0])[i3].T
    assert np.max(np.abs(got - ref)) <= 1e-3


def test_host_entry_staging_from_concurrent_threads(gpu):
    """Four threads call the synchronous host entry at once (ctypes releases
    the GIL), each with its own 40 MiB A, so their pageable buffers go through
    the shared pinned staging buffers together: every result must equal the
    same product computed alone."""
    import threading
    import ttm_amd
    na, pia, q, m = [160, 256, 256], [1, 2, 3], 2, 24
    nc = list(na)
    nc[q - 1] = m
    wa, wc = ttm_amd.dense_strides(na, pia), ttm_amd.dense_strides(nc, pia)
    rng = np.random.default_rng(31)
    inputs = [(rng.uniform(-1, 1, int(np.prod(na))).astype(np.float32),
               rng.uniform(-1, 1, m * na[q - 1]).astype(np.float32)) for _ in range(4)]

    def product(a, b):
        c = np.full(int(np.prod(nc)), np.nan, dtype=np.float32)
        ttm_amd.ttm_raw(q, 3, a, na, wa, pia, b, [m, na[q - 1]], [2, 1], c, nc, wc)
        return c

    alone = [product(a, b) for a, b in inputs]
    together = [None] * len(inputs)
    errors = []

    def work(i):
        try:
            for _ in range(3):
                together[i] = product(*inputs[i])
        except Exception as e:  # reported below
            errors.append(repr(e))

    threads = [threading.Thread(target=work, args=(i,)) for i in range(len(inputs))]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    assert not errors, errors
    assert not any(t.is_alive() for t in threads)
    for i in range(len(inputs)):
        np.testing.assert_array_equal(together[i], alone[i])


@pytest.mark.parametrize("program", ["interfaces", "device_tensor"])
def test_cpp_dropin_program_runs(gpu, program):
    """tests/cpp/interfaces.cpp: the three call forms + error texts, on the GPU.
    tests/cpp/device_tensor.cpp: HBM-resident tensors give the host-tensor results bit for bit."""
    import shutil
    exe = ROOT / "tests" / "_build" / program
    if not exe.exists():
        gxx = shutil.which("g++")
        if not gxx:
            pytest.skip("no g++ to build the program")
        exe.parent.mkdir(exist_ok=True)
        subprocess.run([gxx, "-O1", "-std=c++17", f"-I{ROOT / 'include'}", str(ROOT / "tests" / "cpp" / f"{program}.cpp"),
                        "-o", str(exe), f"-L{ROOT / 'ttm_amd' / 'lib'}", "-lttm_amd",
                        f"-Wl,-rpath,{ROOT / 'ttm_amd' / 'lib'}"], check=True)
    res = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    _say(res.stdout.strip(), res.stderr.strip()[:2000])
    assert res.returncode == 0 and res.stdout.startswith("OK")
