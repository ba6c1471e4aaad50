"""Every global load and store of the tile kernels stays inside its operand:
a CPU model of the address arithmetic of gemm_tile (ttm_gemm.hip: k_gemm in
every compiled configuration, both views, both B orientations, padded slabs,
contraction splits, and the grouped k_gemm_grouped tiles) and lowp_tile
(ttm_lowp.hip: k_lowp with pre-converted B planes, k_lowp_grouped) and the
short-row chain kernels (ttm_rows.hip: k_rows_grouped, k_rows_bf16x3), run for
every thread of a block on the boundary cases -- the minimal interior tile,
edge tiles, the last (ragged) slice of the contraction, ragged m, rows
shorter than a vector.

A GPU memory fault once came from interior B loads of 16- and 32-row tiles
that re-read an index past the tile (the first version's fallback for
vectors past B_VECS): the model runs exactly those loads.  The grouped
configuration list is read from the source and the k_gemm list mirrors
launch_layout's instances.  No GPU is needed.
"""
import itertools
import pathlib
import re

import numpy as np

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
GEMM_SRC = (ROOT / "ttm_amd" / "csrc" / "ttm_gemm.hip").read_text()


class Bounds:
    """Collects every access (operand, first element, element count) and
    checks it against the operand's extent."""

    def __init__(self, sizes):
        self.sizes = sizes
        self.count = 0

    def __call__(self, op, start, n=1):
        self.count += 1
        size = self.sizes[op]
        assert 0 <= start and start + n <= size, (op, start, n, size)


# --------------------------------------------------------------------- gemm
def gemm_configs():
    """(T, JT, NT, WM, BK, TN, BCOL, PAD) of every k_gemm instance launch_layout makes."""
    out = []
    for t, vec in (("float", 4), ("double", 2)):
        base = [(1, 4, 1, 16), (2, 4, 1, 16), (4, 4, 1, 16), (4, 4, 2, 16)]
        if t == "double":
            base += [(2, 2, 1, 16), (4, 2, 1, 16)]
        for tn in (True, False):
            cfgs = list(base)
            if tn:
                cfgs += [(1, 4, 1, 32)] if t == "float" else [(1, 2, 1, 16)]
            for cfg, bcol in itertools.product(cfgs, (True, False)):
                for pad in ((False,) if tn else (False, True)):
                    out.append((t, vec) + cfg + (tn, bcol, pad))
    return out


def group_configs():
    body = GEMM_SRC[GEMM_SRC.index("k_group_configs[] = {"):]
    body = body[:body.index("};")]
    cfgs = [tuple(int(x) for x in m.groups()) for m in re.finditer(r"\{(\d+), (\d+), (\d+), (\d+)\}", body)]
    assert len(cfgs) >= 10
    return cfgs


def gemm_tile_accesses(chk, vec, JT, NT, WM, BK, view, bcol, pad, *, ncols, nq, ld, w, wp, m, j0, n0, flags,
                       a_off, b_off, c_off, a_len):
    """Every global access of one gemm_tile call (ttm_gemm.hip), offsets in
    elements of the whole operands (a_off/b_off/c_off: the split's shift)."""
    TN = view != "NN"
    TSTORE = view == "TN"
    BM = WM * 16 * JT
    BN = (4 // WM) * 16 * NT
    a_vec = bool(flags & 1)
    c_vec = bool(flags & 2)
    b_vec = bool(flags & 8)
    A_VECS = BK * BN // vec
    A_PER = (A_VECS + 255) // 256
    B_VECS = BK * BM // vec
    B_PER = (B_VECS + 255) // 256
    b_vec_ok = b_vec and ((m % vec == 0) if bcol else (ld % vec == 0))
    nslices = (nq + BK - 1) // BK
    threads = []
    for tid in range(256):
        th = {"a": [], "b": []}
        for e in range(A_PER):
            v = tid + 256 * e
            ok = v < A_VECS
            nv = vec
            if not TN:
                row = v // (BN // vec)
                col = (v % (BN // vec)) * vec
                n = n0 + col
                ok = ok and n < ncols
                if pad:
                    k = n // wp if ok else 0
                    i = n - k * wp if ok else 0
                    ok = ok and i < w
                    nv = min(vec, w - i) if ok else 0
                    base = k * ld * w + i if ok else 0
                else:
                    k = n // w if ok else 0
                    base = k * ld * w + (n - k * w if ok else 0)
            else:
                col = v // (BK // vec)
                row = (v % (BK // vec)) * vec
                n = n0 + col
                ok = ok and n < ncols
                base = n * ld if ok else 0
            th["a"].append((v < A_VECS, ok, row, base, nv))
        for e in range(B_PER):
            v = tid + 256 * e
            b_in = v < B_VECS
            if bcol:
                brow, bc = v // (BM // vec), (v % (BM // vec)) * vec
            else:
                bc, brow = v // (BK // vec), (v % (BK // vec)) * vec
            th["b"].append((b_in, brow, bc))
        threads.append(th)

    rows_ok = j0 + BM <= m or not bcol or m % vec == 0
    full = n0 + BN <= ncols and rows_ok and ((a_vec and b_vec_ok) if view == "TT" else (TN or a_vec or not pad))
    if pad and not a_vec:
        full = n0 + BN <= ncols and rows_ok and all(nv == vec for th in threads for _, _, _, _, nv in th["a"])

    def load_full(l0):
        for th in threads:
            for in_a, ok, row, base, nv in th["a"]:
                fix = base + row if TN else base + row * w
                step = 1 if TN else w
                chk("a", a_off + fix + l0 * step, vec)
            for b_in, brow, bc in th["b"]:
                jj = bc if b_in else 0
                ll = brow if b_in else 0
                j = j0 + jj
                if bcol:
                    p = (j if j + vec <= m else m - vec) + (l0 + ll) * m
                else:
                    p = (j if j < m else m - 1) * ld + l0 + ll
                chk("b", b_off + p, vec)

    def load_guarded(l0):
        for th in threads:
            for in_a, ok, row, base, nv in th["a"]:
                if not TN:
                    l = l0 + row
                    if not pad or a_vec:
                        if ok and l < nq:
                            chk("a", a_off + base + l * w, vec)
                    else:
                        lc = l if l < nq else nq - 1
                        addr = base + lc * w
                        lim = a_len - vec
                        ac = addr if addr < lim else lim
                        chk("a", a_off + ac, vec)
                elif nq >= vec:
                    l = l0 + row
                    lc = min(l, nq - vec)
                    chk("a", a_off + base + lc, vec)
                else:
                    l = l0 + row
                    if ok and l < nq:
                        if a_vec:
                            chk("a", a_off + base + l, vec)
                        elif l + vec <= nq:
                            chk("a", a_off + base + l, vec)
                        else:
                            for x in range(vec):
                                if l + x < nq:
                                    chk("a", a_off + base + l + x)
            for b_in, brow, bc in th["b"]:
                if not b_in:
                    continue
                if bcol:
                    l = l0 + brow
                    j = j0 + bc
                    if l < nq:
                        if j + vec <= m:
                            chk("b", b_off + j + l * m, vec)
                        else:
                            for x in range(vec):
                                if j + x < m:
                                    chk("b", b_off + j + x + l * m)
                elif nq >= vec:
                    j = j0 + bc
                    jc = j if j < m else m - 1
                    l = l0 + brow
                    lc = min(l, nq - vec)
                    chk("b", b_off + jc * ld + lc, vec)
                else:
                    j = j0 + bc
                    l = l0 + brow
                    if j < m:
                        if l + vec <= nq:
                            chk("b", b_off + j * ld + l, vec)
                        else:
                            for x in range(vec):
                                if l + x < nq:
                                    chk("b", b_off + j * ld + l + x)

    if view == "TT":
        for s in range(nslices):
            (load_full if (full and nq % BK == 0) else load_guarded)(s * BK)
    else:
        nwhole = nq // BK if full else 0
        for s in range(nslices):
            (load_full if s < nwhole else load_guarded)(s * BK)

    # epilogue
    acc_rows = [4 * lr + r for lr in range(4) for r in range(4)] if vec == 4 else \
               [4 * r + lr for lr in range(4) for r in range(4)]
    for tid in range(256):
        lane, wave = tid & 63, tid >> 6
        lc, lr = lane & 15, lane >> 4
        wm, wn = wave % WM, wave // WM
        a_off_l = wn * 16 * NT + NT * lc
        ncol0 = n0 + a_off_l
        rows = [(4 * lr + r) if vec == 4 else (4 * r + lr) for r in range(4)]
        if not TSTORE:
            for g in range(NT // vec):
                n = ncol0 + g * vec
                inn = n < ncols
                if view == "NN" and not pad:
                    k = n // w if inn else 0
                    gnv = vec if inn else 0
                    cg = k * m * w + (n - k * w if inn else 0)
                else:
                    k = n // wp if inn else 0
                    i = n - k * wp if inn else 0
                    gnv = min(vec, w - i) if (inn and i < w) else 0
                    cg = k * m * w + i
                for jt in range(JT):
                    for ar in rows:
                        j = j0 + wm * 16 * JT + JT * ar + jt
                        if j >= m or gnv == 0:
                            continue
                        if c_vec or (view == "NN" and not pad):
                            chk("c", c_off + cg + j * w, vec)
                        elif view == "NN" and gnv == vec:
                            chk("c", c_off + cg + j * w, vec)
                        else:
                            for x in range(gnv):
                                chk("c", c_off + cg + j * w + x)
        else:
            for nt in range(NT):
                k = ncol0 + nt
                if k >= ncols:
                    continue
                for ar in rows:
                    jr = j0 + wm * 16 * JT + JT * ar
                    if JT % vec == 0 and c_vec and jr + JT <= m:
                        for jt in range(0, JT, vec):
                            chk("c", c_off + k * m + jr + jt, vec)
                        continue
                    for jt in range(JT):
                        if jr + jt < m:
                            chk("c", c_off + k * m + jr + jt)


def gemm_problems(vec, tn, bm, bn, bk, pad):
    """(outer, nq, w, m) boundary problems for one tile shape."""
    # ragged m with m % vec == 0 (vec, bm + vec, 2 bm - vec) runs the unguarded loop
    ms = sorted({1, 3, vec, bm - 1, bm, bm + 1, bm + vec, 2 * bm - vec, 2 * bm + 5} - {0})
    nqs = sorted({1, vec - 1, bk, bk + 1, 2 * bk + 3, 64 + 5} - {0})
    if tn:
        shapes = [(outer, 1) for outer in (1, 5, bn, bn + 3)]
    elif pad:
        shapes = [(outer, w) for outer, w in ((1, 3), (2, 2), (3, bn - 1), (2, bn + 3), (1, 5), (4, 7))]
    else:
        shapes = [(outer, w) for outer, w in ((1, bn), (2, vec), (3, bn + vec), (1, 3 * vec))]
    for (outer, w), nq, m in itertools.product(shapes, nqs, ms):
        if outer * nq * w >= vec:
            yield outer, nq, w, m


@pytest.mark.parametrize("cfg", gemm_configs(), ids=lambda c: "-".join(map(str, c)))
def test_k_gemm_accesses_in_bounds(cfg):
    t, vec, JT, NT, WM, BK, tn, bcol, pad = cfg
    BM, BN = WM * 16 * JT, (4 // WM) * 16 * NT
    checked = 0
    for outer, nq, w, m in gemm_problems(vec, tn, BM, BN, BK, pad):
        wp = 1 if tn else (w + vec - 1) // vec * vec
        ncols = outer if tn else outer * wp
        sizes = {"a": outer * nq * w, "b": m * nq, "c": outer * m * w}
        nj, nn = (m + BM - 1) // BM, (ncols + BN - 1) // BN
        # the launcher's flags for 16-byte aligned buffers (pad: A's vectors
        # element-aligned; the padded instance runs when w % vec != 0)
        a_vec = (nq % vec == 0) if tn else (w % vec == 0 and not pad)
        c_vec = (m % vec == 0) if tn else (w % vec == 0 and not pad)
        flags = (1 if a_vec else 0) | (2 if c_vec else 0) | 8
        # one split, and a contraction cut into ranges of 64 when long enough
        splits = [(0, nq)]
        if nq > 64:
            splits = [(l0, min(64, nq - l0)) for l0 in range(0, nq, 64)]
        tiles = {(nj - 1, nn - 1), (0, 0), (max(0, nj - 2), max(0, nn - 2))}
        for (l0, ln), (jb, nb) in itertools.product(splits, tiles):
            c_off = 0
            chk = Bounds(dict(sizes, c=sizes["c"] if len(splits) == 1 else sizes["c"] * len(splits)))
            if len(splits) > 1:
                c_off = (l0 // 64) * sizes["c"]
            a_off = l0 * (1 if tn else w)
            b_off = l0 * (m if bcol else 1)
            a_len = 0 if tn else (ncols // wp) * nq * w - l0 * w
            gemm_tile_accesses(chk, vec, JT, NT, WM, BK, "TN" if tn else "NN", bcol, pad, ncols=ncols, nq=ln,
                               ld=nq, w=w, wp=wp, m=m, j0=jb * BM, n0=nb * BN, flags=flags, a_off=a_off,
                               b_off=b_off, c_off=c_off, a_len=a_len)
            checked += chk.count
    assert checked > 0


@pytest.mark.parametrize("gcfg", group_configs(), ids=lambda c: "-".join(map(str, c)))
@pytest.mark.parametrize("vec", [4, 2])
def test_k_gemm_grouped_accesses_in_bounds(gcfg, vec):
    JT, NT, WM, BK = gcfg
    if vec == 2 and BK == 32:
        pytest.skip("slices of 32 are compiled for float32 only")
    BM, BN = WM * 16 * JT, (4 // WM) * 16 * NT
    for ncols, nq, m in itertools.product((1, 3, BN - 1, BN, BN + 5), (1, vec - 1, BK, BK + 3, 40, 256),
                                          (1, 4, BM - 1, BM, BM + 7)):
        if nq == 0:
            continue
        sizes = {"a": ncols * nq, "b": m * nq, "c": m * ncols}
        a_al, c_al = True, True
        flags = (1 if (a_al and nq % vec == 0) else 0) | (2 if (c_al and ncols % vec == 0) else 0) | 8
        nj, nn = (m + BM - 1) // BM, (ncols + BN - 1) // BN
        for jb, nb in {(nj - 1, nn - 1), (0, 0), (0, nn - 1), (nj - 1, 0)}:
            chk = Bounds(sizes)
            gemm_tile_accesses(chk, vec, JT, NT, WM, BK, "TT", False, False, ncols=ncols, nq=nq, ld=nq, w=ncols,
                               wp=ncols, m=m, j0=jb * BM, n0=nb * BN, flags=flags, a_off=0, b_off=0, c_off=0,
                               a_len=0)


# --------------------------------------------------------------------- lowp
def lds_row_inv(P):
    return (P & ~63) | ((P & 15) << 2) | ((P & 63) >> 4)


def lowp_tile_accesses(chk, view, bsplit, BM, *, ncols, nq, w, m, j0, n0, a_vec, ldb=0, plane=0, parts=1):
    """Global loads of one lowp_tile call (ttm_lowp.hip), in float elements of
    A and B, and 16-byte pieces of the converted planes."""
    BN, BK = 128, 32
    BH = BM // 64
    TN = view != "NN"
    full = n0 + BN <= ncols and nq % BK == 0 and (not TN or a_vec)
    nslices = (nq + BK - 1) // BK
    for tid in range(256):
        lane, wave = tid & 63, tid >> 6
        mn_half, mn_y, mn_h = lane & 1, (lane >> 1) & 15, lane >> 5
        mn_c4 = 64 * mn_h + 4 * mn_y
        mn_k4 = 8 * wave + 4 * mn_half
        k_c, k_p = tid & 3, tid >> 2
        k_row = lds_row_inv(k_p)
        a_base, a_ok = [0, 0], [False, False]
        if not TN:
            n = n0 + mn_c4
            a_ok[0] = n < ncols
            k = n // w if a_ok[0] else 0
            a_base[0] = k * nq * w + (n - k * w if a_ok[0] else 0) + mn_k4 * w
        else:
            for s in range(2):
                n = n0 + k_row + 64 * s
                a_ok[s] = n < ncols
                a_base[s] = (n * nq if a_ok[s] else 0) + 8 * k_c
        for sl in range(nslices):
            l0 = sl * BK
            if full:
                if not TN:
                    for e in range(4):
                        chk("a", a_base[0] + l0 * w + e * w, 4)
                else:
                    for e in range(4):
                        chk("a", a_base[e >> 1] + l0 + 4 * (e & 1), 4)
            else:
                if not TN:
                    for e in range(4):
                        l = l0 + mn_k4 + e
                        if a_ok[0] and l < nq:
                            chk("a", a_base[0] + (l0 + e) * w, 4)
                else:
                    for e in range(4):
                        s = e >> 1
                        l = l0 + 8 * k_c + 4 * (e & 1)
                        if a_ok[s] and l < nq:
                            if a_vec or l + 4 <= nq:
                                chk("a", a_base[s] + l0 + 4 * (e & 1), 4)
                            else:
                                for x in range(4):
                                    if l + x < nq:
                                        chk("a", a_base[s] + l0 + 4 * (e & 1) + x)
            if bsplit:
                for p in range(parts):
                    for s in range(BH):
                        # i32x4 = 8 two-byte elements of a plane
                        chk("bs", p * plane + (j0 + k_row + 64 * s) * ldb + l0 + 8 * k_c, 8)
            else:
                for e in range(2 * BH):
                    j = j0 + k_row + 64 * (e >> 1)
                    l = l0 + 8 * k_c + 4 * (e & 1)
                    if j >= m:
                        continue
                    if l + 4 <= nq:
                        chk("b", j * nq + l, 4)
                    else:
                        for x in range(4):
                            if l + x < nq:
                                chk("b", j * nq + l + x)


@pytest.mark.parametrize("BM", [64, 128])
@pytest.mark.parametrize("view", ["NN", "TN"])
@pytest.mark.parametrize("parts", [1, 2, 3])
def test_k_lowp_accesses_in_bounds(BM, view, parts):
    """k_lowp<MODE, TN, BM>: A from the product, B from its planes (rows
    padded to whole blocks, ldb = n_q rounded up to 32, as k_lowp_split_b
    writes them)."""
    vec = 4
    for nq, m in itertools.product((1, 3, 32, 33, 64, 100), (1, 5, BM - 1, BM, BM + 3)):
        shapes = [(1, 1), (3, 1), (129, 1), (300, 1)] if view == "TN" else [(1, 128), (2, 4), (3, 132), (1, 256)]
        for outer, w in shapes:
            ncols = outer if view == "TN" else outer * w
            nj = (m + BM - 1) // BM
            ldb = (nq + 31) // 32 * 32
            rows = nj * BM
            plane = rows * ldb
            a_vec = (nq % vec == 0) if view == "TN" else True
            sizes = {"a": outer * nq * w, "b": m * nq, "bs": parts * plane}
            nn = (ncols + 127) // 128
            for jb, nb in {(nj - 1, nn - 1), (0, 0)}:
                chk = Bounds(sizes)
                lowp_tile_accesses(chk, view, True, BM, ncols=ncols, nq=nq, w=w, m=m, j0=jb * BM, n0=nb * 128,
                                   a_vec=a_vec, ldb=ldb, plane=plane, parts=parts)


def lds_swz(P):
    return ((P >> 1) & 3) ^ ((P >> 3) & 1)


def lowp_dma_b_accesses(chk, BM, *, j0, nq, ldb, plane, parts):
    """B's planes by LDS DMA in the pipelined BF16X3 tile (dma_b of
    ttm_lowp.hip), for every slice the loop requests (slices past the end
    repeat the last one), with each lane's LDS destination checked against
    its part of the tile: lane L of the 1 KB piece at physical row p0 writes
    physical row p0 + L / 4, position L % 4."""
    BK = 32
    BH = BM // 64
    nslices = (nq + BK - 1) // BK
    last = (nslices - 1) * BK
    for t in range(nslices + 2):
        l0 = min(t * BK, last)
        for wv in range(4):
            for p in range(parts):
                for h in range(BH):
                    p0 = 64 * h + 16 * wv
                    for lane in range(64):
                        pr = p0 + (lane >> 2)
                        assert pr < BM and (lane & 3) < 4
                        off = (lds_row_inv(pr) & 63) * ldb + 8 * ((lane & 3) ^ lds_swz(pr))
                        # the logical row the register path stores at physical row pr
                        assert lds_row_inv(pr) // 64 == h
                        chk("bs", p * plane + (j0 + 64 * h) * ldb + l0 + off, 8)


@pytest.mark.parametrize("view", ["NN", "TN"])
def test_k_lowp_pipelined_accesses_in_bounds(view):
    """k_lowp<2, TN, 128> (BF16X3, pipelined): A's loads as in the other
    tiles (two slices ahead, clamped to the last slice), B's planes by LDS
    DMA; the planes are padded to whole blocks and slices."""
    BM, parts = 128, 3
    for nq, m in itertools.product((4, 32, 33, 100, 256), (65, 128, 131, 300)):
        shapes = [(1, 1), (3, 1), (129, 1), (300, 1)] if view == "TN" else [(1, 128), (2, 4), (3, 132), (1, 256)]
        for outer, w in shapes:
            ncols = outer if view == "TN" else outer * w
            nj = (m + BM - 1) // BM
            ldb = (nq + 31) // 32 * 32
            plane = nj * BM * ldb
            sizes = {"a": outer * nq * w, "b": m * nq, "bs": parts * plane}
            nn = (ncols + 127) // 128
            for jb, nb in {(nj - 1, nn - 1), (0, 0)}:
                chk = Bounds(sizes)
                lowp_tile_accesses(chk, view, True, BM, ncols=ncols, nq=nq, w=w, m=m, j0=jb * BM, n0=nb * 128,
                                   a_vec=(nq % 4 == 0) if view == "TN" else True, ldb=ldb, plane=plane,
                                   parts=parts)
                lowp_dma_b_accesses(chk, BM, j0=jb * BM, nq=nq, ldb=ldb, plane=plane, parts=parts)


def test_dma_lane_order_matches_the_register_layout():
    """Each (physical row, position) of a B part is written by exactly one
    DMA lane, with the logical row and chunk the register path puts there
    (lds_at(P, c) = P * 32 + 8 * (c ^ lds_swz(P)))."""
    BM = 128
    seen = {}
    for wv in range(4):
        for h in range(2):
            p0 = 64 * h + 16 * wv
            for lane in range(64):
                pr, pos = p0 + (lane >> 2), lane & 3
                c = pos ^ lds_swz(pr)
                seen[(pr, pos)] = (lds_row_inv(pr), c)
    assert len(seen) == BM * 4
    for (pr, pos), (row, c) in seen.items():
        # the register path: thread (k_p, k_c) holds logical row lds_row_inv(k_p) (+64), chunk k_c,
        # and stores it at lds_at(k_p + 64 s, k_c)
        assert pr * 32 + 8 * pos == pr * 32 + 8 * (c ^ lds_swz(pr))
        assert row == lds_row_inv(pr)


@pytest.mark.parametrize("BM", [64, 128])
def test_k_lowp_grouped_accesses_in_bounds(BM):
    """k_lowp_grouped: TT tiles converting A and the row-major B themselves."""
    for ncols, nq, m in itertools.product((1, 5, 127, 128, 131), (1, 3, 32, 33, 100), (1, 4, BM - 1, BM, BM + 2)):
        sizes = {"a": ncols * nq, "b": m * nq}
        nj, nn = (m + BM - 1) // BM, (ncols + 127) // 128
        for jb, nb in {(nj - 1, nn - 1), (0, 0)}:
            chk = Bounds(sizes)
            lowp_tile_accesses(chk, "TT", False, BM, ncols=ncols, nq=nq, w=ncols, m=m, j0=jb * BM, n0=nb * 128,
                               a_vec=nq % 4 == 0)


def test_the_model_catches_the_fault_it_was_written_for():
    """The first version of the interior B loads sent vectors past B_VECS
    (BM * BK < 256 * VEC) to the thread's first index b_col = tid / (BK /
    VEC), which is not a row of the tile: the model must flag that on the
    16-row tile (a planted mistake in the model, not in the kernel)."""
    vec, JT, NT, WM, BK = 4, 1, 4, 1, 16
    BM = WM * 16 * JT
    m, nq = 16, 64
    sizes = {"b": m * nq}
    chk = Bounds(sizes)
    B_VECS = BK * BM // vec
    with pytest.raises(AssertionError):
        for tid in range(256):
            b_in = tid < B_VECS
            bc, brow = tid // (BK // vec), (tid % (BK // vec)) * vec
            jj = bc  # the old fallback kept the thread's own column for b_in == False
            ll = brow if b_in else 0
            chk("b", (0 + jj) * nq + (nq - BK) + ll, vec)


def test_transpose_grouped_tiles_cover_every_element_once():
    """k_transpose_grouped (chain stages over modes without a factor): the
    tile enumeration of the stage (nj = ceil(nq / 64) l-blocks per 64-column
    block, l-blocks fastest, problems concatenated) reads only inside each
    A, writes only inside each C, and writes every C element exactly once
    with A's element at the transposed place."""
    rng = np.random.default_rng(5)
    probs = [(int(rng.integers(1, 200)), int(rng.integers(1, 150))) for _ in range(12)] + [(64, 64), (1, 1), (65, 63)]
    tiles = []
    for pid, (ncols, nq) in enumerate(probs):
        nj, nn = (nq + 63) // 64, (ncols + 63) // 64
        begin = len(tiles)
        tiles += [(pid, begin, nj)] * (nj * nn)
    srcs = [rng.standard_normal(ncols * nq) for ncols, nq in probs]
    outs = [np.full(ncols * nq, np.nan) for ncols, nq in probs]
    writes = [np.zeros(ncols * nq, dtype=int) for ncols, nq in probs]
    for t, (pid, begin, nj) in enumerate(tiles):
        ncols, nq = probs[pid]
        local = t - begin
        l0, n0 = (local % nj) * 64, (local // nj) * 64
        tile = np.zeros((64, 65))
        for ty in range(4):
            for tx in range(64):
                for r in range(ty, 64, 4):
                    n, l = n0 + r, l0 + tx
                    if n < ncols and l < nq:
                        assert 0 <= n * nq + l < ncols * nq
                        tile[r][tx] = srcs[pid][n * nq + l]
        for ty in range(4):
            for tx in range(64):
                for r in range(ty, 64, 4):
                    l, n = l0 + r, n0 + tx
                    if l < nq and n < ncols:
                        assert 0 <= l * ncols + n < ncols * nq
                        outs[pid][l * ncols + n] = tile[tx][r]
                        writes[pid][l * ncols + n] += 1
    for (ncols, nq), a, c, w in zip(probs, srcs, outs, writes):
        assert (w == 1).all()
        assert np.array_equal(c.reshape(nq, ncols), a.reshape(ncols, nq).T)


def rows_unit(chk, A, B, nq, m, ncols, u, MB):
    """k_rows_grouped (ttm_rows.hip), one unit of 16 columns: every global
    load of A and B and every LDS read of the staged B, and the unit's
    16 x MB block of C from the model of the lanes' fragments and the
    16x16x4 MFMA (D[i][c] += sum_q src0(lane i + 16 q) * src1(lane c + 16 q))."""
    S = (nq + 15) // 16 * 16 + 4
    ng = (nq + 15) // 16
    width = 16 * ng
    total = ncols * nq
    # B staged: rows past m and columns past nq zero
    sB = np.full(MB * S, np.nan)
    for idx in range(MB * width):
        j, l = idx // width, idx % width
        assert j * S + l < MB * (256 + 4)
        if j < m and l < nq:
            chk("b", j * nq + l)
            sB[j * S + l] = B[j * nq + l]
        else:
            sB[j * S + l] = 0.0
    col0 = u * 16
    D = np.zeros((MB, 16))
    for lane in range(64):
        lc, lq = lane % 16, lane // 16
        n = col0 + lc if col0 + lc < ncols else ncols - 1
        frags = []
        for g in range(ng - 1):
            s = n * nq + 4 * lq + 16 * g
            chk("a", s, 4)
            assert 16 * g + 4 * lq + 3 < nq  # inside row n
            frags.append(A[s:s + 4])
        lim = nq - (16 * (ng - 1) + 4 * lq)
        e = n * nq + 16 * (ng - 1) + 4 * lq
        ec = e if e + 4 <= total else total - 4
        d = e - ec
        chk("a", ec, 4)
        rl = A[ec:ec + 4]
        af = np.zeros(4)
        for x in range(4):
            t = rl[x] if d == 0 else (rl[x + d] if x + d < 4 else 0.0)
            af[x] = t if x < lim else 0.0
            if x < lim:
                assert x + d < 4 and ec + x + d == e + x
        frags.append(af)
        for g, fr in enumerate(frags):
            for x in range(4):
                for i in range(MB):
                    addr = (i % 16 + 16 * (i // 16)) * S + 4 * lq + 16 * g + x
                    assert 0 <= addr < MB * S and 16 * g + 4 * lq + x < width
                    D[i, lc] += sB[addr] * fr[x]
    for lc in range(16):
        if col0 + lc < ncols:
            for j in range(m):
                chk("c", j * ncols + col0 + lc)
    return D


@pytest.mark.parametrize("MB", [16, 32, 48, 64])
def test_k_rows_grouped_unit_model(MB):
    """The short-row kernel's lane arithmetic on the CPU: loads stay inside
    A (the last group of the last row clamps to A's end and shifts), the
    staged B covers every fragment read, and the MFMA model of a unit
    equals B A^T for contraction lengths around every multiple of 4 and 16,
    the extremes 4 and 256, partial last units and a single column."""
    rng = np.random.default_rng(MB)
    cases = [(4, 1), (5, 3), (7, 17), (15, 16), (16, 33), (17, 40), (31, 2), (33, 47), (100, 5), (255, 18),
             (256, 16), (129, 1)]
    for nq, ncols in cases:
        m = int(rng.integers(MB // 2 + 1, MB + 1)) if MB > 16 else int(rng.integers(1, 17))
        A = rng.standard_normal(ncols * nq)
        B = rng.standard_normal(m * nq)
        chk = Bounds({"a": ncols * nq, "b": m * nq, "c": m * ncols})
        want = B.reshape(m, nq) @ A.reshape(ncols, nq).T
        for u in range((ncols + 15) // 16):
            D = rows_unit(chk, A, B, nq, m, ncols, u, MB)
            cols = min(16, ncols - 16 * u)
            np.testing.assert_allclose(D[:m, :cols], want[:, 16 * u:16 * u + cols], rtol=1e-12, atol=1e-12)


def rows_unit_bf16x3(chk, A, B, nq, m, ncols, u, MB):
    """k_rows_bf16x3 (ttm_rows.hip), one unit: groups of 32 contraction
    indices, lane (c, q) reads two 16-byte pieces A[n][32 g + 8 q ..] (the
    last group's pieces each clamped to A's end and shifted), and the staged
    B planes (row stride 32 k + 8 bf16) are read at lc * S + 8 q + 32 g
    (+ 16 jt rows).  The split itself is exact in sum (x0 + x1 + x2 = x up to
    2^-24 |x|), so the model multiplies the unsplit values."""
    S = (nq + 31) // 32 * 32 + 8
    ng = (nq + 31) // 32
    width = 32 * ng
    total = ncols * nq
    sB = np.zeros(MB * S)
    for j in range(MB):
        for l in range(0, width, 2):
            for x in (0, 1):
                if j < m and l + x < nq:
                    chk("b", j * nq + l + x)
                    sB[j * S + l + x] = B[j * nq + l + x]
            assert j * S + l + 1 < MB * (256 + 8)
    col0 = u * 16
    D = np.zeros((MB, 16))

    def last_piece(e, lim):
        ec = e if e + 4 <= total else total - 4
        d = e - ec
        chk("a", ec, 4)
        rl = A[ec:ec + 4]
        out = np.zeros(4)
        for x in range(4):
            if x < lim:
                assert x + d < 4 and ec + x + d == e + x
                out[x] = rl[x + d]
        return out

    for lane in range(64):
        lc, lq = lane % 16, lane // 16
        n = col0 + lc if col0 + lc < ncols else ncols - 1
        groups = []
        for g in range(ng - 1):
            s = n * nq + 32 * g + 8 * lq
            chk("a", s, 4)
            chk("a", s + 4, 4)
            assert 32 * g + 8 * lq + 7 < nq
            groups.append(np.concatenate([A[s:s + 4], A[s + 4:s + 8]]))
        e0 = n * nq + 32 * (ng - 1) + 8 * lq
        lim0 = nq - (32 * (ng - 1) + 8 * lq)
        groups.append(np.concatenate([last_piece(e0, lim0), last_piece(e0 + 4, lim0 - 4)]))
        for g, fr in enumerate(groups):
            for e in range(8):
                for i in range(MB):
                    addr = i * S + 8 * lq + 32 * g + e
                    assert 0 <= addr < MB * S and 32 * g + 8 * lq + e < width
                    D[i, lc] += sB[addr] * fr[e]
    for lc in range(16):
        if col0 + lc < ncols:
            for j in range(m):
                chk("c", j * ncols + col0 + lc)
    return D


@pytest.mark.parametrize("MB", [16, 32, 48, 64])
def test_k_rows_bf16x3_unit_model(MB):
    """The BF16X3 short-row kernel's lane arithmetic (32-index groups, two
    clamped pieces in the last group, three B planes) on the CPU: in-bounds
    loads and the unit's block of C equal to B A^T.  k_rows_fp16x2 takes the
    same groups, pieces and plane layout (two planes)."""
    rng = np.random.default_rng(100 + MB)
    cases = [(4, 1), (5, 3), (31, 17), (32, 16), (33, 33), (63, 40), (100, 5), (255, 18), (256, 16), (129, 1)]
    for nq, ncols in cases:
        m = int(rng.integers(MB // 2 + 1, MB + 1)) if MB > 16 else int(rng.integers(1, 17))
        A = rng.standard_normal(ncols * nq)
        B = rng.standard_normal(m * nq)
        chk = Bounds({"a": ncols * nq, "b": m * nq, "c": m * ncols})
        want = B.reshape(m, nq) @ A.reshape(ncols, nq).T
        for u in range((ncols + 15) // 16):
            D = rows_unit_bf16x3(chk, A, B, nq, m, ncols, u, MB)
            cols = min(16, ncols - 16 * u)
            np.testing.assert_allclose(D[:m, :cols], want[:, 16 * u:16 * u + cols], rtol=1e-12, atol=1e-12)


def rows_unit_f64(chk, A, B, nq, m, ncols, u, MB):
    """k_rows_f64 (ttm_rows.hip), one unit: groups of 8 contraction indices,
    lane (c, q) reads the 2-element piece A[n][8 g + 2 q ..] (the last
    group's clamped to A's last two elements and shifted), B staged with row
    stride 8 k + 2 and read at (16 jt + c) S + 2 q + 8 g, accumulator rows
    16 jt + 4 r + q."""
    S = (nq + 7) // 8 * 8 + 2
    ng = (nq + 7) // 8
    width = 8 * ng
    total = ncols * nq
    sB = np.full(MB * S, np.nan)
    for j in range(MB):
        for l in range(width):
            assert j * S + l < MB * (256 + 2)
            if j < m and l < nq:
                chk("b", j * nq + l)
                sB[j * S + l] = B[j * nq + l]
            else:
                sB[j * S + l] = 0.0
    col0 = u * 16
    D = np.zeros((MB, 16))
    for lane in range(64):
        lc, lq = lane % 16, lane // 16
        n = col0 + lc if col0 + lc < ncols else ncols - 1
        groups = []
        for g in range(ng - 1):
            s = n * nq + 2 * lq + 8 * g
            chk("a", s, 2)
            assert 8 * g + 2 * lq + 1 < nq
            groups.append(A[s:s + 2])
        lim = nq - (8 * (ng - 1) + 2 * lq)
        e = n * nq + 8 * (ng - 1) + 2 * lq
        ec = e if e + 2 <= total else total - 2
        d = e - ec
        chk("a", ec, 2)
        rl = A[ec:ec + 2]
        af = np.zeros(2)
        for x in range(2):
            if x < lim:
                assert x + d < 2 and ec + x + d == e + x
                af[x] = rl[x + d]
        groups.append(af)
        for g, fr in enumerate(groups):
            for x in range(2):
                for i in range(MB):
                    addr = i * S + 2 * lq + 8 * g + x
                    assert 0 <= addr < MB * S and 8 * g + 2 * lq + x < width
                    D[i, lc] += sB[addr] * fr[x]
    for lc in range(16):
        if col0 + lc < ncols:
            for r in range(4):
                for jt in range(MB // 16):
                    for q in range(4):  # lane quarter q holds rows 16 jt + 4 r + q
                        jj = 16 * jt + 4 * r + q
                        if jj < m:
                            chk("c", jj * ncols + col0 + lc)
    return D


@pytest.mark.parametrize("MB", [16, 32, 48, 64])
def test_k_rows_f64_unit_model(MB):
    """The float64 short-row kernel's lane arithmetic on the CPU (8-index
    groups, a clamped 2-element last piece, B stride 8 k + 2): in-bounds
    loads and the unit's block of C equal to B A^T."""
    rng = np.random.default_rng(200 + MB)
    cases = [(4, 1), (5, 3), (7, 17), (8, 16), (9, 33), (15, 40), (100, 5), (255, 18), (256, 16), (129, 1)]
    for nq, ncols in cases:
        m = int(rng.integers(MB // 2 + 1, MB + 1)) if MB > 16 else int(rng.integers(1, 17))
        A = rng.standard_normal(ncols * nq)
        B = rng.standard_normal(m * nq)
        chk = Bounds({"a": ncols * nq, "b": m * nq, "c": m * ncols})
        want = B.reshape(m, nq) @ A.reshape(ncols, nq).T
        for u in range((ncols + 15) // 16):
            D = rows_unit_f64(chk, A, B, nq, m, ncols, u, MB)
            cols = min(16, ncols - 16 * u)
            np.testing.assert_allclose(D[:m, :cols], want[:, 16 * u:16 * u + cols], rtol=1e-12, atol=1e-12)
