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
    for prec in ("bf16", "fp16", "bf16x3", "fp8"):
        np.testing.assert_array_equal(le.emulate(prec, A3, B), exact, err_msg=prec)
