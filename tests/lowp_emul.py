"""CPU emulation of the reduced-precision arithmetics (test infrastructure).

Written from the format definitions, independent of the GPU code:
* bfloat16: the upper 16 bits of binary32 with round-to-nearest-even;
* binary16: numpy's float16 conversion (round-to-nearest-even);
* BF16X3: x = x0 + x1 + x2 with x_k = bf16(x - x0 - .. - x_{k-1}) and the six
  products of total order <= 2 (Henry, Tang & Heinecke, ARITH 2019);
* FP16X2: each operand scaled by a power of two 2^(15 - E), where f * 2^E
  (f in [0.5, 1)) is the largest finite magnitude it is taken over (none
  non-zero: 2^0), then x = x0 + x1 with x0 = fp16(x), x1 = fp16(x - x0),
  and the three products x0 y0 + x0 y1 + x1 y0, scaled back by the powers of
  two.  B takes one scale; A one scale in total, or (the tiled kernel's
  64-row blocks) one per block of 128 columns and slice of 32 contraction
  indices: the running minimum of the exponents of the block's slices so far;
* OCP FP8 e4m3fn: 1 sign, 4 exponent (bias 7), 3 mantissa bits, no
  infinities, max 448, subnormal step 2^-9; round-to-nearest-even; with one
  power-of-two scale 2^e per block of 32 consecutive contraction indices,
  e = ceil(log2(amax / 448)).
Each function returns float64 arrays holding the exactly representable values.
"""
from __future__ import annotations

import numpy as np


def bf16(x):
    x = np.ascontiguousarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    rounded = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    out = rounded.astype(np.uint32).view(np.float32).astype(np.float64)
    return np.where(np.isnan(x), np.nan, out)


def fp16(x):
    return np.asarray(x, dtype=np.float32).astype(np.float16).astype(np.float64)


def bf16x3_parts(x):
    x = np.asarray(x, dtype=np.float32)
    p0 = bf16(x)
    r1 = (x.astype(np.float64) - p0).astype(np.float32)  # exact in float32
    p1 = bf16(r1)
    r2 = (r1.astype(np.float64) - p1).astype(np.float32)
    p2 = bf16(r2)
    return p0, p1, p2


def fp16x2_exponent(x):
    """15 - E for the largest finite magnitude f * 2^E of x (0 if none is non-zero)."""
    a = np.abs(np.asarray(x, dtype=np.float64)).reshape(-1)
    a = a[np.isfinite(a)]
    amax = float(a.max()) if a.size else 0.0
    if not (amax > 0.0):
        return 0
    return 15 - int(np.frexp(amax)[1])


def fp16x2_parts(x, e):
    """Two binary16 parts of x * 2^e (x float32, the scaling exact)."""
    xs = (np.asarray(x, dtype=np.float64) * 2.0 ** e).astype(np.float32)
    p0 = fp16(xs)
    r1 = (xs.astype(np.float64) - p0).astype(np.float32)  # exact in float32
    return p0, fp16(r1)


def e4m3fn(v):
    """Round float64 values (|v| <= 448 after scaling) to e4m3fn, RNE."""
    v = np.asarray(v, dtype=np.float64)
    a = np.abs(v)
    e = np.floor(np.log2(np.where(a > 0, a, 1.0)))
    e = np.maximum(e, -6.0)                       # subnormals share the 2^-6 binade step
    step = np.exp2(e - 3.0)
    q = np.round(a / step) * step                 # numpy rounds half to even
    q = np.minimum(q, 448.0)
    return np.sign(v) * q


def _finite_abs(blk):
    """|x| of finite values, 0 for infinities and NaNs (they set no scale)."""
    a = np.abs(blk)
    return np.where(np.isfinite(a), a, 0.0)


def fp8_block_quantize(x, axis):
    """Quantize `x` (float32) along `axis` in blocks of 32 with E8M0 scales.
    Infinities and NaNs set no scale and quantize to NaN (e4m3fn has no
    infinity)."""
    x = np.asarray(x, dtype=np.float64)
    x = np.moveaxis(x, axis, -1)
    n = x.shape[-1]
    out = np.empty_like(x)
    for b0 in range(0, n, 32):
        blk = x[..., b0:b0 + 32]
        amax = np.max(_finite_abs(blk), axis=-1, keepdims=True)
        with np.errstate(divide="ignore"):
            e = np.where(amax > 0, np.ceil(np.log2(np.where(amax > 0, amax, 1.0) / 448.0)), 0.0)
        e = np.where(amax * np.exp2(-e) > 448.0, e + 1, e)
        e = np.clip(e, -126, 127)
        fin = np.isfinite(blk)
        with np.errstate(invalid="ignore"):
            q = e4m3fn(np.clip(np.where(fin, blk, 0.0) * np.exp2(-e), -448.0, 448.0)) * np.exp2(e)
        out[..., b0:b0 + 32] = np.where(fin, q, np.nan)
    return np.moveaxis(out, -1, axis)


def fp8_block_quantum(x, axis):
    """The e4m3fn subnormal step 2^(e-9) of each element's 32-block (its
    smallest non-zero magnitude after scaling back)."""
    x = np.asarray(x, dtype=np.float64)
    x = np.moveaxis(x, axis, -1)
    out = np.empty_like(x)
    for b0 in range(0, x.shape[-1], 32):
        blk = x[..., b0:b0 + 32]
        amax = np.max(_finite_abs(blk), axis=-1, keepdims=True)
        with np.errstate(divide="ignore"):
            e = np.where(amax > 0, np.ceil(np.log2(np.where(amax > 0, amax, 1.0) / 448.0)), 0.0)
        e = np.where(amax * np.exp2(-e) > 448.0, e + 1, e)
        out[..., b0:b0 + 32] = np.exp2(np.clip(e, -126, 127) - 9) * np.ones_like(blk)
    return np.moveaxis(out, -1, axis)


def _fp16x2_sliced(A3, Bm, cols=128, bk=32):
    """FP16X2 with A scaled per block of `cols` fused columns n = k*w + i and
    slice of `bk` contraction indices (running minimum of the exponents)."""
    outer, nq, w = A3.shape
    eb = fp16x2_exponent(Bm)
    b0, b1 = fp16x2_parts(Bm, eb)
    acol = A3.transpose(0, 2, 1).reshape(outer * w, nq)
    out = np.zeros((outer * w, Bm.shape[0]))
    for c0 in range(0, outer * w, cols):
        ea = None
        for l0 in range(0, nq, bk):
            blk = acol[c0:c0 + cols, l0:l0 + bk]
            fin = np.abs(blk.astype(np.float64))
            fin = fin[np.isfinite(fin)]
            if fin.size and fin.max() > 0:
                e = 15 - int(np.frexp(float(fin.max()))[1])
                ea = e if ea is None else min(ea, e)
            es = 0 if ea is None else ea
            a0, a1 = fp16x2_parts(blk, es)
            p0, p1 = b0[:, l0:l0 + bk], b1[:, l0:l0 + bk]
            out[c0:c0 + cols] += (a0 @ p0.T + a1 @ p0.T + a0 @ p1.T) * 2.0 ** -(es + eb)
    return out.reshape(outer, w, -1).transpose(0, 2, 1)


def emulate(prec, A3, Bm, sliced=False):
    """C3[k][j][i] = sum_l B(j,l) A3[k][l][i] with the operands as `prec`
    rounds them, products and sums in float64 (reference for the GPU
    result up to FP32 accumulation error).  A3: (outer, nq, w), Bm: (m, nq).
    `sliced` selects FP16X2's per-block and per-slice scale of A (the tiled
    kernel's 64-row blocks) instead of one scale for all of A."""
    A3 = np.asarray(A3, dtype=np.float32)
    Bm = np.asarray(Bm, dtype=np.float32)
    if prec == "bf16":
        return np.einsum("jl,kli->kji", bf16(Bm), bf16(A3))
    if prec == "fp16":
        return np.einsum("jl,kli->kji", fp16(Bm), fp16(A3))
    if prec == "bf16x3":
        a = bf16x3_parts(A3)
        b = bf16x3_parts(Bm)
        total = 0
        for p, q in ((0, 0), (0, 1), (1, 0), (0, 2), (1, 1), (2, 0)):
            total = total + np.einsum("jl,kli->kji", b[p], a[q])
        return total
    if prec == "fp16x2" and sliced:
        return _fp16x2_sliced(A3, Bm)
    if prec == "fp16x2":
        ea, eb = fp16x2_exponent(A3), fp16x2_exponent(Bm)
        a = fp16x2_parts(A3, ea)
        b = fp16x2_parts(Bm, eb)
        total = 0
        for p, q in ((0, 0), (0, 1), (1, 0)):
            total = total + np.einsum("jl,kli->kji", b[p], a[q])
        return total * 2.0 ** -(ea + eb)
    if prec == "fp8":
        return np.einsum("jl,kli->kji", fp8_block_quantize(Bm, axis=1), fp8_block_quantize(A3, axis=1))
    raise ValueError(prec)
