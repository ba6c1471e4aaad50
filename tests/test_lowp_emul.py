"""CPU checks of the reduced-precision emulation that the GPU variants are compared with."""
import numpy as np

import lowp_emul as le


def test_bf16_rounding_known_answers():
    x = np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -9, 1 + 2 ** -7 + 2 ** -8, -2.5, 3.0e38], dtype=np.float32)
    got = le.bf16(x)
    expect = [1.0, 1.0, 1 + 2 ** -7, 1 + 2 ** -6, -2.5, 1.765625 * 2.0 ** 127]
    np.testing.assert_allclose(got, expect, rtol=1e-7)


def test_bf16x3_split_is_exact_to_24_bits():
    rng = np.random.default_rng(0)
    x = rng.uniform(-1, 1, 10000).astype(np.float32)
    p0, p1, p2 = le.bf16x3_parts(x)
    rel = np.abs(p0 + p1 + p2 - x.astype(np.float64)) / np.abs(x.astype(np.float64))
    assert rel.max() < 2.0 ** -24
    # dropping the third part is an error of order 2^-17 -- the split matters
    assert (np.abs(p0 + p1 - x) / np.abs(x)).max() > 2.0 ** -20


def test_fp16x2_split_and_scale():
    """The scale puts the largest magnitude in [2^14, 2^15); two binary16 parts
    then hold 22 significant bits, so dropping x1 y1 leaves about 2^-22."""
    assert le.fp16x2_exponent(np.array([0.75, -0.1])) == 15 and le.fp16x2_exponent(np.array([1.0])) == 14
    assert le.fp16x2_exponent(np.array([3.0e38])) == 15 - 128 and le.fp16x2_exponent(np.zeros(3)) == 0
    assert le.fp16x2_exponent(np.array([1.0, np.inf, np.nan])) == 14
    rng = np.random.default_rng(3)
    for mag in (1e-30, 1e-3, 1.0, 1e5, 1e30):
        x = (rng.uniform(-1, 1, 10000) * mag).astype(np.float32)
        e = le.fp16x2_exponent(x)
        assert 2.0 ** 14 <= np.abs(x).max() * 2.0 ** e < 2.0 ** 15
        p0, p1 = le.fp16x2_parts(x, e)
        back = (p0 + p1) * 2.0 ** -e
        big = np.abs(x) > np.abs(x).max() * 2.0 ** -8        # x1 above the binary16 subnormal range
        rel = np.abs(back - x.astype(np.float64))[big] / np.abs(x.astype(np.float64))[big]
        assert rel.max() <= 2.0 ** -22
        # every element: 2^-22 relative, or half the binary16 subnormal step
        # (2^-25 scaled back) where the second part is subnormal
        assert np.all(np.abs(back - x.astype(np.float64)) <= 2.0 ** -22 * np.abs(x.astype(np.float64))
                      + 2.0 ** -25 * 2.0 ** -e)
        # one part alone is only binary16 accurate: the second part matters
        assert (np.abs(p0 * 2.0 ** -e - x)[big] / np.abs(x)[big]).max() > 2.0 ** -13


def test_fp16x2_product_error_is_fp32_level():
    """FP16X2 against the exact float64 product: normwise error below 1e-6 on
    uniform and wide-exponent data (the three dropped-term and rounding terms
    are each below 2^-22 relative)."""
    rng = np.random.default_rng(4)
    for wide in (False, True):
        A3 = rng.uniform(-1, 1, (4, 256, 6))
        B = rng.uniform(-1, 1, (9, 256))
        if wide:
            A3 = A3 * np.exp2(rng.integers(-30, 30, A3.shape))
        A3, B = A3.astype(np.float32), B.astype(np.float32)
        exact = np.einsum("jl,kli->kji", B.astype(np.float64), A3.astype(np.float64))
        for sliced in (False, True):
            got = le.emulate("fp16x2", A3, B, sliced=sliced)
            err = np.linalg.norm(got - exact) / (np.linalg.norm(A3.astype(np.float64)) *
                                                 np.linalg.norm(B.astype(np.float64)))
            assert err < 1e-6, (wide, sliced, err)


def test_fp16x2_slice_scales_follow_blocks_and_slices():
    """The per-slice rule: a block whose later slice holds a larger value
    lowers its exponent from that slice on (earlier partial sums rescaled
    exactly), other blocks keep theirs; the result equals the float64
    product for data the parts represent exactly (small integers times
    powers of two), where one scale for all of A loses a tiny value."""
    rng = np.random.default_rng(5)
    A3 = rng.integers(-8, 9, (1, 96, 300)).astype(np.float32)   # 300 columns: blocks of 128, 128, 44
    A3[0, 70, 5] = 2.0 ** 20           # block 0, slice 2: its exponent drops there
    A3[0, 10, 200] = 2.0 ** -30        # block 1: kept by the binary16 parts of its own scale
    B = rng.integers(-8, 9, (5, 96)).astype(np.float32)
    exact = np.einsum("jl,kli->kji", B.astype(np.float64), A3.astype(np.float64))
    np.testing.assert_array_equal(le.emulate("fp16x2", A3, B, sliced=True), exact)
    assert not np.array_equal(le.emulate("fp16x2", A3, B), exact)


def test_e4m3fn_known_answers():
    # values read back from the gfx950 conversion instruction (in-bounds probe):
    # 1.0 -> 0x38, 448 -> 0x7e, -1.5 -> 0xbc, 0.001 -> 0x01, 0.0625 -> 0x18, 3e-3 -> 0x02
    got = le.e4m3fn([1.0, 448.0, -1.5, 0.001, 0.0, 0.0625, 3.0e-3])
    expect = [1.0, 448.0, -1.5, 2.0 ** -9, 0.0, 0.0625, 2 * 2.0 ** -9]
    np.testing.assert_array_equal(got, expect)
    # every e4m3fn value is a fixed point; midpoints round to even mantissas
    assert le.e4m3fn(1 + 2 ** -4) == 1.0 and le.e4m3fn(1 + 3 * 2 ** -4) == 1.25


def test_fp8_block_scaling_bounds_error():
    rng = np.random.default_rng(1)
    x = rng.uniform(-1, 1, (7, 96)).astype(np.float32) * np.exp2(rng.integers(-20, 20, (7, 1)))
    q = le.fp8_block_quantize(x, axis=1)
    for b0 in range(0, 96, 32):
        blk, qb = x[:, b0:b0 + 32].astype(np.float64), q[:, b0:b0 + 32]
        amax = np.abs(blk).max(axis=1, keepdims=True)
        # relative to the block's scale the rounding error is at most half a step
        assert np.all(np.abs(qb - blk) <= amax * 2.0 ** -3 + 1e-300)


def test_emulation_matches_float64_for_exact_inputs():
    """Small integers are exact in every format, so every emulation equals the product."""
    rng = np.random.default_rng(2)
    A3 = rng.integers(-8, 9, (3, 40, 5)).astype(np.float32)
    B = rng.integers(-8, 9, (6, 40)).astype(np.float32)
    exact = np.einsum("jl,kli->kji", B.astype(np.float64), A3.astype(np.float64))
    for prec in ("bf16", "fp16", "bf16x3", "fp8", "fp16x2"):
        np.testing.assert_array_equal(le.emulate(prec, A3, B), exact, err_msg=prec)
    np.testing.assert_array_equal(le.emulate("fp16x2", A3, B, sliced=True), exact)


def test_fp8_non_finite_values_set_no_scale():
    """A block with an infinity and a NaN keeps the scale of its finite values
    (the same quantized values as with those entries zeroed) and quantizes
    both non-finite entries to NaN; other blocks are unaffected."""
    rng = np.random.default_rng(3)
    x = rng.uniform(-5, 5, (3, 96)).astype(np.float32)
    y = x.copy()
    y[1, 40] = np.inf
    y[1, 41] = np.nan
    y[2, 90] = -np.inf
    qx = le.fp8_block_quantize(np.where(np.isfinite(y), y, 0), axis=1)
    qy = le.fp8_block_quantize(y, axis=1)
    bad = ~np.isfinite(y)
    assert np.isnan(qy[bad]).all()
    assert np.array_equal(qy[~bad], qx[~bad])
    # the scale is the finite maximum's: a value at the block maximum keeps
    # e4m3fn precision (relative step 2^-3), not flushed by a huge scale
    blk = y[1, 32:64]
    j = int(np.argmax(np.where(np.isfinite(blk), np.abs(blk), 0)))
    assert abs(qy[1, 32 + j] - blk[j]) <= 2.0 ** -3 * abs(blk[j])
