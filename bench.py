#!/usr/bin/env python3
"""Benchmark of the mode-q TTM on one MI355X.

Workload (one "step"): the plain FP32/FP64 drop-in at scale -- an order-3
tensor A of extents 1024 x 1024 x 512 in float32 and in float64, first-order
(1,2,3) and last-order (3,2,1) layouts, contraction mode q in {1,2,3}, B of
shape m x n_q with m in {16, 1024}, entries uniform[-1,1]: 24 products, all
inputs resident in HBM before timing starts.

Printed JSON line:
  value      aggregate TFLOP/s over the step (algorithmic 2*m*N_A flops)
  roofline   the dominant kernel (largest share of step time): algorithmic
             work / its launch time against the MI355X peak for its bound
  cpu_baseline  the oracle's restatement of the reference algorithm (OpenMP
             loop over BLAS GEMMs, `combined` dispatch) timed on the host on a
             bounded sample of the same workload
  cases      per-product time, TFLOP/s, GB/s and roofline fraction
"""
from __future__ import annotations

import argparse
import json
import os
import pathlib
import sys
import time

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

PEAK_HBM = 8.0e12                      # B/s
PEAK_FLOPS = {"f32": 157.3e12, "f64": 78.6e12, "bf16": 2516.6e12}
EXTENTS = [1024, 1024, 512]
LAYOUTS = {"first": [1, 2, 3], "last": [3, 2, 1]}
MS = [16, 1024]
METRIC = "TTM TFLOP/s + HBM GB/s vs roofline; BF16x3 vs FP32 at equal FP64-oracle error"

# HBM traffic per launch from two rocprofv3 --pmc passes (FETCH_SIZE, then
# WRITE_SIZE; KiB per dispatch, mean over the dispatches of the case), keyed by
# kernel symbol and case.  Both counters are scaled by what they read for a
# plain 1 GiB device copy (`bench.py --copy-calibration`): FETCH_SIZE reads
# 524,303 KiB for 1,048,576 KiB fetched (scale 2.0), WRITE_SIZE 1,048,576 KiB
# for 1,048,576 KiB written (scale 1.0).  Summaries and commands:
# profiles/ttm_hbm_traffic_pmc.md.  A symbol or case without an entry
# reports traffic null.
PMC_FETCH_SCALE = 2.0
PMC_WRITE_SCALE = 1.0
PMC_TRAFFIC_KIB = {
    "k_gemm<double, 4, 4, 2, 16, false, false, false>": {"f64-first-q2-m1024": (6264054.0, 4200535.7)},
    "k_gemm<float, 4, 4, 2, 16, false, false, false>": {"f32-first-q2-m1024": (3081001.4, 2097152.0)},
    "k_nn_stream<float, 1, 6, true>": {"f32-first-q2-m16": (1048885.0, 32768.0),
                                        "f32-first-q3-m16": (1048757.6, 65536.0)},
    "k_tn_lds<float, 8, 2, 257>": {"f32-first-q1-m16": (1053808.5, 32768.0)},
    "k_tn_lds<double, 3, 2, 513>": {"f64-first-q1-m16": (2099106.4, 65537.5)},
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--dtypes", default="f32,f64")
    ap.add_argument("--only", default="", help="comma list of case labels to run (profiling)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="CPU baseline budget (0 = skip)")
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--force", default="", help="force a kernel family (comparison runs only)")
    ap.add_argument("--no-bf16x3", dest="bf16x3", action="store_false",
                    help="skip the BF16X3-vs-FP32 comparison on the compute-bound float32 products")
    ap.add_argument("--tucker-count", type=int, default=256,
                    help="tensors of the batched uneven Tucker-chain section (0 = skip)")
    ap.add_argument("--no-host-path", dest="host_path", action="store_false",
                    help="skip the host-pointer (PCIe-inclusive) drop-in measurement")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device copy (scale for FETCH_SIZE / WRITE_SIZE counters)")
    ap.add_argument("--no-extras", dest="extras", action="store_false",
                    help="skip the configuration-1 latency, order-5, column-major-B, split and odd-extent sections")
    return ap.parse_args()


def copy_calibration():
    """Three launches of a plain device-to-device copy of 1 GiB (1 GiB read,
    1 GiB written per launch), the known-size kernel against which the
    FETCH_SIZE / WRITE_SIZE counter readings are scaled to HBM bytes."""
    import torch
    n = 1 << 28
    src = torch.ones(n, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_read": 4 * n, "copy_calibration_bytes_written": 4 * n,
                      "launches": 3}), flush=True)


# An untimed read of 512 MiB before every timed launch evicts the 256 MiB
# Infinity Cache (and the L2s), so that no product reads its inputs from a
# cache the previous one filled.  A read, not a write: the write-back
# Infinity Cache would otherwise still hold up to 256 MiB of dirty lines
# when the product starts, and their write-back competes with the product's
# reads (measured on a 1.1 GB GEMV: 169 us unflushed, 247 us after a 512 MiB
# write, 173 us after a 512 MiB read).
FLUSH_BYTES = 512 << 20


class Flusher:
    def __init__(self, torch, dev):
        self.buf = torch.ones(FLUSH_BYTES // 4, dtype=torch.float32, device=dev)
        self.sink = None

    def __call__(self):
        self.sink = self.buf.sum()


def timed_ms(torch, run, steps, stream, flush=None, warmup=2):
    """Mean and minimum HIP-event time of `run` on `stream` over `steps`
    launches, each after an untimed cache flush."""
    for _ in range(warmup):
        run()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    for e0, e1 in ev:
        if flush is not None:
            flush()
        e0.record(stream)
        run()
        e1.record(stream)
    torch.cuda.synchronize()
    ms = [e0.elapsed_time(e1) for e0, e1 in ev]
    return float(np.mean(ms)), float(np.min(ms))


def roof(flops, bytes_, ms, dtype):
    """Bound and fraction of the roofline of one product."""
    t = ms * 1e-3
    t_hbm = bytes_ / PEAK_HBM
    t_cmp = flops / PEAK_FLOPS[dtype]
    return {"bound": "hbm" if t_hbm >= t_cmp else "mfma", "frac": round(max(t_hbm, t_cmp) / t, 4),
            "gbps": round(bytes_ / t / 1e9, 1), "tflops": round(flops / t / 1e12, 3)}


def case_list(dtypes):
    out = []
    for dt in dtypes:
        for lay, pia in LAYOUTS.items():
            for q in (1, 2, 3):
                for m in MS:
                    out.append(dict(dtype=dt, layout=lay, pia=pia, q=q, m=m,
                                    label=f"{dt}-{lay}-q{q}-m{m}"))
    return out


def work(na, q, m, itemsize):
    n_a = int(np.prod(na))
    n_c = n_a // na[q - 1] * m
    flops = 2.0 * m * n_a
    bytes_ = itemsize * (n_a + m * na[q - 1] + n_c)
    return flops, bytes_


def gpu_bench(args, cases):
    import torch
    import ttm_amd

    dev = torch.device("cuda:0")
    tdt = {"f32": torch.float32, "f64": torch.float64}
    n_a = int(np.prod(EXTENTS))
    gen = torch.Generator(device=dev)
    gen.manual_seed(0x77140002)
    A = {}
    Cbuf = {}
    for dt in {c["dtype"] for c in cases}:
        A[dt] = torch.empty(n_a, dtype=tdt[dt], device=dev).uniform_(-1, 1, generator=gen)
        max_c = max(n_a // EXTENTS[c["q"] - 1] * c["m"] for c in cases if c["dtype"] == dt)
        Cbuf[dt] = torch.empty(max_c, dtype=tdt[dt], device=dev)
    for c in cases:
        nq = EXTENTS[c["q"] - 1]
        c["B"] = torch.empty((c["m"], nq), dtype=tdt[c["dtype"]], device=dev).uniform_(-1, 1, generator=gen)
        c["na"] = list(EXTENTS)
        c["wa"] = ttm_amd.dense_strides(EXTENTS, c["pia"])
        c["nc"] = list(EXTENTS)
        c["nc"][c["q"] - 1] = c["m"]
        c["wc"] = ttm_amd.dense_strides(c["nc"], c["pia"])
        c["kernel"] = ttm_amd.plan_kernel(c["q"], EXTENTS, c["pia"], [c["m"], nq], [2, 1],
                                          is_double=c["dtype"] == "f64")
        c["symbol"] = ttm_amd.plan_kernel_symbol(c["q"], EXTENTS, c["pia"], [c["m"], nq], [2, 1],
                                                 is_double=c["dtype"] == "f64")
        c["flops"], c["bytes"] = work(EXTENTS, c["q"], c["m"], 4 if c["dtype"] == "f32" else 8)
    torch.cuda.synchronize()

    def run(c):
        ttm_amd.ttm_raw(c["q"], 3, A[c["dtype"]], c["na"], c["wa"], c["pia"], c["B"], [c["m"], EXTENTS[c["q"] - 1]],
                        [2, 1], Cbuf[c["dtype"]], c["nc"], c["wc"])

    for _ in range(args.warmup):
        for c in cases:
            run(c)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    flush = Flusher(torch, dev)
    events = [[(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in cases]
              for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(args.steps):
        for c, (e0, e1) in zip(cases, events[s]):
            flush()
            e0.record(stream)
            run(c)
            e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    for i, c in enumerate(cases):
        ms = [events[s][i][0].elapsed_time(events[s][i][1]) for s in range(args.steps)]
        c["ms"] = float(np.mean(ms))
        c["ms_min"] = float(np.min(ms))
        t = c["ms"] * 1e-3
        c["tflops"] = c["flops"] / t / 1e12
        # the reference's count, m * (N_A / n_q) * (2 n_q - 1) (example/measure.cpp:70-73)
        nq = EXTENTS[c["q"] - 1]
        c["tflops_reference_count"] = c["m"] * (c["flops"] / (2.0 * c["m"] * nq)) * (2 * nq - 1) / t / 1e12
        c["gbps"] = c["bytes"] / t / 1e9
        t_hbm = c["bytes"] / PEAK_HBM
        t_cmp = c["flops"] / PEAK_FLOPS[c["dtype"]]
        c["bound"] = "hbm" if t_hbm >= t_cmp else "mfma"
        c["frac"] = max(t_hbm, t_cmp) / t
    split = None
    if args.bf16x3 and "f32" in A:
        split = bf16x3_vs_fp32(args, cases, A["f32"], Cbuf["f32"], stream)
    colb = colmajor_m16(args, cases, A, Cbuf, stream, flush) if args.extras else None
    del A, Cbuf, flush
    torch.cuda.empty_cache()
    return wall / args.steps, split, colb


def colmajor_m16(args, cases, A, Cbuf, stream, flush):
    """The m = 16 products of the sweep again with B column-major (pib = {1,2},
    the layout of the reference's tensor object, include/tlib/detail/tensor.h:84-87)."""
    import torch
    import ttm_amd
    rows = []
    for c in cases:
        if c["m"] != 16:
            continue
        nq = EXTENTS[c["q"] - 1]
        bcol = c["B"].t().contiguous()  # n_q x m storage = B column-major
        dbl = c["dtype"] == "f64"
        run = lambda: ttm_amd.ttm_raw(c["q"], 3, A[c["dtype"]], c["na"], c["wa"], c["pia"], bcol, [c["m"], nq],
                                      [1, 2], Cbuf[c["dtype"]], c["nc"], c["wc"])
        ms, ms_min = timed_ms(torch, run, args.steps, stream, flush)
        r = roof(c["flops"], c["bytes"], ms, c["dtype"])
        rows.append({"label": c["label"] + "-bcol",
                     "symbol": ttm_amd.plan_kernel_symbol(c["q"], EXTENTS, c["pia"], [c["m"], nq], [1, 2], is_double=dbl),
                     "ms": round(ms, 4), "ms_min": round(ms_min, 4), **r,
                     "row_major_frac": round(c["frac"], 4)})
        del bcol
    return rows


OTHER_PRECISIONS = ("bf16", "fp16", "fp8")


def bf16x3_vs_fp32(args, cases, A, C, stream):
    """The compute-bound float32 products again in BF16X3 and FP16X2 (and, for
    reference, plain BF16, FP16 and scaled FP8-e4m3): time, and the normwise
    relative error of each arithmetic against the float64 definition on 4096
    sampled entries of C."""
    import torch
    import ttm_amd
    rng = np.random.default_rng(17)
    rows = []
    for c in cases:
        if c["dtype"] != "f32" or c["m"] != 1024:
            continue
        q = c["q"]
        nq = EXTENTS[q - 1]

        def run(prec):
            ttm_amd.ttm_raw(q, 3, A, c["na"], c["wa"], c["pia"], c["B"], [c["m"], nq], [2, 1], C, c["nc"], c["wc"],
                            precision=prec)

        res = {}
        for prec in ("default", "bf16x3", "fp16x2") + OTHER_PRECISIONS:
            for _ in range(2):
                run(prec)
            ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
            for e0, e1 in ev:
                e0.record(stream)
                run(prec)
                e1.record(stream)
            torch.cuda.synchronize()
            ms = float(np.mean([e0.elapsed_time(e1) for e0, e1 in ev]))
            # sampled error against the float64 definition
            idx = [rng.integers(0, c["nc"][r], size=4096) for r in range(3)]
            offc = sum(idx[r] * c["wc"][r] for r in range(3))
            base = sum(idx[r] * c["wa"][r] for r in range(3) if r != q - 1)
            fib = torch.from_numpy((base[:, None] + np.arange(nq)[None, :] * c["wa"][q - 1]).astype(np.int64))
            a_f = A[fib.to(A.device)].double().cpu().numpy()
            b_r = c["B"].double().cpu().numpy()[idx[q - 1]]
            expect = np.sum(a_f * b_r, axis=1)
            got = C[torch.from_numpy(offc.astype(np.int64)).to(A.device)].double().cpu().numpy()
            res[prec] = (ms, float(np.linalg.norm(got - expect) / np.linalg.norm(expect)))
        rows.append({"case": c["label"], "fp32_ms": round(res["default"][0], 4),
                     "bf16x3_ms": round(res["bf16x3"][0], 4),
                     "speedup": round(res["default"][0] / res["bf16x3"][0], 3),
                     "fp32_normwise_err": res["default"][1], "bf16x3_normwise_err": res["bf16x3"][1],
                     "bf16x3_fp32equiv_tflops": round(c["flops"] / res["bf16x3"][0] / 1e9, 2),
                     # six bf16 products per FP32-equivalent product
                     "bf16x3_mfma_issued_tflops": round(6 * c["flops"] / res["bf16x3"][0] / 1e9, 2),
                     "fp16x2_ms": round(res["fp16x2"][0], 4),
                     "fp16x2_speedup": round(res["default"][0] / res["fp16x2"][0], 3),
                     "fp16x2_normwise_err": res["fp16x2"][1],
                     "fp16x2_fp32equiv_tflops": round(c["flops"] / res["fp16x2"][0] / 1e9, 2),
                     # three fp16 products per FP32-equivalent product (plus the scale pass over A)
                     "fp16x2_mfma_issued_tflops": round(3 * c["flops"] / res["fp16x2"][0] / 1e9, 2),
                     # the plain single-part arithmetics (errors at their format's level)
                     "other": {pr: {"ms": round(res[pr][0], 4), "speedup": round(res["default"][0] / res[pr][0], 3),
                                    "normwise_err": res[pr][1]} for pr in OTHER_PRECISIONS}})
    if not rows:
        return None
    sp = [r["speedup"] for r in rows]
    sp2 = [r["fp16x2_speedup"] for r in rows]
    return {"speedup_geomean": round(float(np.exp(np.mean(np.log(sp)))), 3),
            "fp16x2_speedup_geomean": round(float(np.exp(np.mean(np.log(sp2)))), 3),
            "max_fp16x2_normwise_err": max(r["fp16x2_normwise_err"] for r in rows),
            "max_bf16x3_normwise_err": max(r["bf16x3_normwise_err"] for r in rows),
            "max_fp32_normwise_err": max(r["fp32_normwise_err"] for r in rows),
            "other_speedup_geomean": {pr: round(float(np.exp(np.mean(np.log([r["other"][pr]["speedup"] for r in rows])))), 3)
                                      for pr in OTHER_PRECISIONS},
            "other_max_normwise_err": {pr: max(r["other"][pr]["normwise_err"] for r in rows) for pr in OTHER_PRECISIONS},
            "error_reference": "float64 definition on 4096 sampled entries", "cases": rows}


def tucker_items(torch, dev, count, seed=0x77140005):
    """The Tucker batch: `count` first-order float32 tensors with extents
    drawn from U{16..256}, factors with orthonormal rows (QR of a Gaussian
    matrix) and ranks drawn from U{4..64}.  Returns the chain items
    (A, [U1, U2, U3], C), the ranks, the algorithmic flops and the compulsory
    bytes."""
    rng = np.random.default_rng(seed)
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    items = []
    ranks = []
    flops = 0.0
    compulsory = 0.0
    for _ in range(count):
        na = [int(x) for x in rng.integers(16, 257, size=3)]
        # a factor with orthonormal rows needs r <= n: ranks drawn above the
        # extent are clamped to it
        rk = [min(int(x), n) for x, n in zip(rng.integers(4, 65, size=3), na)]
        # A dense in first-order layout: natural shape, strides (1, n1, n1*n2)
        A = torch.empty(na[::-1], device=dev, dtype=torch.float32).uniform_(-1, 1, generator=gen).permute(2, 1, 0)
        Us = []
        for r, n in zip(rk, na):
            Q, _ = np.linalg.qr(rng.standard_normal((n, r)))
            Us.append(torch.from_numpy(np.ascontiguousarray(Q[:, :r].T)).to(dev, torch.float32))
        C = torch.empty(rk[::-1], device=dev, dtype=torch.float32).permute(2, 1, 0)
        items.append((A, Us, C))
        ranks.append(rk)
        cur = list(na)
        for k in range(3):  # stages in layout order 1, 2, 3
            flops += 2.0 * rk[k] * float(np.prod(cur))
            cur[k] = rk[k]
        compulsory += 4.0 * (np.prod(na) + sum(r * n for r, n in zip(rk, na)) + np.prod(rk))
    return items, ranks, flops, compulsory


def tucker_launch_roofline(ranks, extents, esz, peak, cus):
    """Sum over the launches of a chain plan (stage s of every tensor, split by
    rank class m <= 16 / 32 / 48 / 64 / larger; ranks 33..48 share the 64-row
    launch in stages where the ranks 33..64 have fewer than 64 units of 16
    columns per compute unit, and all ranks up to 64 share one launch in
    stages where they have fewer than 128 units per compute unit) of each
    launch's own bound: max(HBM bytes / 8 TB/s, matrix flops / peak), the
    ranks padded to the launch's row height and the contraction to whole
    groups of 16 (the short-row kernels' and the tiles' padding).  The plan
    cannot beat this sum (its launches run one after the other within a
    stage and stage by stage)."""
    total = 0.0
    for s in range(3):
        probs = []
        for rk, na in zip(ranks, extents):
            cur = list(na)
            for k in range(s):
                cur[k] = rk[k]
            nq, m = na[s], rk[s]
            probs.append((nq, m, int(np.prod(cur)) // nq))
        split48 = sum((nc + 15) // 16 for _, m, nc in probs if 32 < m <= 64) >= 64 * cus
        small = sum((nc + 15) // 16 for _, m, nc in probs if m <= 64) < 128 * cus
        top = max((m for _, m, _ in probs if m <= 64), default=0)
        top_h = 16 if top <= 16 else 32 if top <= 32 else 48 if top <= 48 else 64
        cls = {}
        for nq, m, ncols in probs:
            if m <= 64 and small:
                h = top_h
            else:
                h = (16 if m <= 16 else 32 if m <= 32 else 48 if m <= 48 and split48 else 64 if m <= 64
                     else 128 * ((m + 127) // 128))
            b, f = cls.get(h, (0.0, 0.0))
            cls[h] = (b + esz * (ncols * nq + ncols * m + m * nq),
                      f + 2.0 * h * (16 * ((nq + 15) // 16)) * ncols)
        total += sum(max(b / PEAK_HBM, f / peak) for b, f in cls.values())
    return total * 1e3  # ms


def tucker_batch(args):
    """Tucker batch (tucker_items) of 256 tensors as one plan, replayed as one
    hipGraph, in FP32, BF16X3 and FP16X2, and the same tensors and factors in
    float64 (FP64: twice the compulsory bytes)."""
    import torch
    from ttm_amd.chain import ChainPlan
    dev = torch.device("cuda:0")
    items, ranks, flops, compulsory = tucker_items(torch, dev, args.tucker_count)
    extents = [list(A.shape) for A, _, _ in items]
    out = {"workload": f"{args.tucker_count} tensors, extents U{{16..256}}^3, ranks U{{4..64}}, "
                       "A x1 U1 x2 U2 x3 U3, orthonormal rows, one hipGraph replay per step",
           "gflop": round(flops / 1e9, 2), "compulsory_gb": round(compulsory / 1e9, 3)}
    stream = torch.cuda.current_stream()
    items64 = [(A.double(), [U.double() for U in Us], C.double()) for A, Us, C in items]
    for prec in ("fp32", "bf16x3", "fp16x2", "fp64"):
        plan = ChainPlan(items64 if prec == "fp64" else items, precision=prec if prec in ("bf16x3", "fp16x2") else None)
        if prec in ("fp32", "fp64"):
            plan.tune()
        plan.run_graph()
        plan.run_graph()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(args.steps):
            plan.run_graph()
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / args.steps
        cb = compulsory * (2 if prec == "fp64" else 1)
        out[prec] = {"ms": round(ms, 3), "tflops": round(flops / ms / 1e9, 3),
                     "compulsory_gbps": round(cb / ms / 1e6, 1),
                     "hbm_frac_of_compulsory": round(cb / ms / 1e6 / (PEAK_HBM / 1e9), 4),
                     "stage_kernels": plan.stage_kernels}
        if prec in ("fp32", "fp64"):
            roof = tucker_launch_roofline(ranks, extents, 8 if prec == "fp64" else 4,
                                          PEAK_FLOPS["f64" if prec == "fp64" else "f32"],
                                          torch.cuda.get_device_properties(dev).multi_processor_count)
            out[prec]["launch_roofline_ms"] = round(roof, 3)
            out[prec]["frac_of_launch_rooflines"] = round(roof / ms, 4)
        plan.close()
    del items, items64
    torch.cuda.empty_cache()
    return out


def _sampled_err(torch, A, na, wa, Bm, C, nc, wc, q, rng, samples=4096):
    """Normwise relative error of sampled entries of C against the float64
    definition (Bm is the m x n_q matrix)."""
    p = len(na)
    idx = [rng.integers(0, nc[r], size=samples) for r in range(p)]
    offc = sum(idx[r] * wc[r] for r in range(p))
    base = sum(idx[r] * wa[r] for r in range(p) if r != q - 1)
    nq = na[q - 1]
    fib = torch.from_numpy((base[:, None] + np.arange(nq)[None, :] * wa[q - 1]).astype(np.int64))
    a_f = A[fib.to(A.device)].double().cpu().numpy()
    b_r = Bm.double().cpu().numpy()[idx[q - 1]]
    expect = np.sum(a_f * b_r, axis=1)
    got = C[torch.from_numpy(offc.astype(np.int64)).to(A.device)].double().cpu().numpy()
    return float(np.linalg.norm(got - expect) / np.linalg.norm(expect))


def _product(torch, dev, na, pia, q, m, colmajor, dtype, gen):
    """Device buffers and arguments of one product with uniform[-1,1] data."""
    import ttm_amd
    tdt = torch.float32 if dtype == "f32" else torch.float64
    nq = na[q - 1]
    A = torch.empty(int(np.prod(na)), dtype=tdt, device=dev).uniform_(-1, 1, generator=gen)
    Bm = torch.empty((m, nq), dtype=tdt, device=dev).uniform_(-1, 1, generator=gen)
    B = Bm.t().contiguous() if colmajor else Bm
    nc = list(na)
    nc[q - 1] = m
    C = torch.empty(int(np.prod(nc)), dtype=tdt, device=dev)
    wa, wc = ttm_amd.dense_strides(na, pia), ttm_amd.dense_strides(nc, pia)
    args = (q, len(na), A, na, wa, pia, B, [m, nq], [1, 2] if colmajor else [2, 1], C, nc, wc)
    return A, Bm, C, args


def _timed_product(torch, args_, na, pia, q, m, colmajor, dtype, gen, flush, precision=None, err=False, rng=None):
    import ttm_amd
    A, Bm, C, pargs = _product(torch, torch.device("cuda:0"), na, pia, q, m, colmajor, dtype, gen)
    stream = torch.cuda.current_stream()
    run = lambda: ttm_amd.ttm_raw(*pargs, precision=precision)
    ms, ms_min = timed_ms(torch, run, args_.steps, stream, flush)
    sym, splits = ttm_amd.last_kernel()
    flops, bytes_ = work(na, q, m, 4 if dtype == "f32" else 8)
    row = {"na": na, "pi": pia, "q": q, "m": m, "b": "col" if colmajor else "row", "dtype": dtype,
           "precision": precision or "default", "symbol": sym, "splits": splits, "ms": round(ms, 4),
           "ms_min": round(ms_min, 4), **roof(flops, bytes_, ms, dtype)}
    if err:
        row["normwise_err_sampled"] = _sampled_err(torch, A, na, pargs[4], Bm, C, pargs[10], pargs[11], q, rng)
    del A, Bm, C, pargs
    return row


def config1_latency(args):
    """A 64x48x32 first-order, q = 2, B 40x48 in both orientations, float32
    and float64: microseconds per call (HIP events around one launch, warm and
    after a cache flush; and the host clock around 200 back-to-back calls of
    the device entry) and the normwise error against the float64 definition."""
    import torch
    import ttm_amd
    dev = torch.device("cuda:0")
    flush = Flusher(torch, dev)
    stream = torch.cuda.current_stream()
    na, pia, q, m = [64, 48, 32], [1, 2, 3], 2, 40
    rows = []
    for dtype, npdt in (("f32", np.float32), ("f64", np.float64)):
        rng = np.random.default_rng(0x77140001)
        a = rng.uniform(-1, 1, int(np.prod(na))).astype(npdt)
        Bm = rng.uniform(-1, 1, (m, na[q - 1])).astype(npdt)
        A3 = a.reshape(na[2], na[1], na[0]).astype(np.float64)
        ref = np.einsum("kli,jl->kji", A3, Bm.astype(np.float64))
        for colmajor in (True, False):
            A = torch.from_numpy(a).to(dev)
            B = torch.from_numpy(np.ascontiguousarray(Bm.T if colmajor else Bm).reshape(-1)).to(dev)
            nc = [64, m, 32]
            C = torch.empty(int(np.prod(nc)), dtype=A.dtype, device=dev)
            wa, wc = ttm_amd.dense_strides(na, pia), ttm_amd.dense_strides(nc, pia)
            pargs = (q, 3, A, na, wa, pia, B, [m, na[q - 1]], [1, 2] if colmajor else [2, 1], C, nc, wc)
            run = lambda: ttm_amd.ttm_raw(*pargs)
            warm, warm_min = timed_ms(torch, run, 200, stream, None, warmup=20)
            cold, _ = timed_ms(torch, run, 50, stream, flush, warmup=2)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(200):
                run()
            torch.cuda.synchronize()
            host_us = (time.perf_counter() - t0) / 200 * 1e6
            got = C.double().cpu().numpy().reshape(32, m, 64)
            den = np.linalg.norm(a.astype(np.float64)) * np.linalg.norm(Bm.astype(np.float64))
            rows.append({"dtype": dtype, "b": "col" if colmajor else "row", "symbol": ttm_amd.last_kernel()[0],
                         "device_us_warm": round(warm * 1e3, 2), "device_us_warm_min": round(warm_min * 1e3, 2),
                         "device_us_after_flush": round(cold * 1e3, 2), "host_us_per_call": round(host_us, 2),
                         "normwise_err": float(np.linalg.norm(got - ref) / den),
                         "normwise_bound_4nq_eps": 4 * na[q - 1] * float(np.finfo(npdt).eps)})
    return rows


def config4_order5(args):
    """Order-5 float32 A 32x24x40x16x48, q = 1..5, B 64 x n_q: FP32 and
    BF16X3, each product after a cache flush (HBM-bound), in two layouts: a
    seeded random permutation (drawn again while it is the identity) and the
    first-order layout."""
    import torch
    dev = torch.device("cuda:0")
    flush = Flusher(torch, dev)
    gen = torch.Generator(device=dev)
    gen.manual_seed(0x77140004)
    rng = np.random.default_rng(0x77140004)
    na = [32, 24, 40, 16, 48]
    first = [1, 2, 3, 4, 5]
    pia = first
    while pia == first:
        pia = [int(x) + 1 for x in rng.permutation(5)]
    rows = []
    for layout in (pia, first):
        for q in range(1, 6):
            for prec in (None, "bf16x3"):
                rows.append(_timed_product(torch, args, na, layout, q, 64, False, "f32", gen, flush,
                                           precision=prec, err=True, rng=rng))
    return {"layouts": [pia, first], "rows": rows,
            "fp32_min_frac": min(r["frac"] for r in rows if r["precision"] == "default"),
            "bf16x3_min_frac": min(r["frac"] for r in rows if r["precision"] == "bf16x3")}


def mid_m_products(args):
    """A 1024x1024x512 first-order with 32, 48 and 64 rows of B (between the
    sweep's 16 and 1024): the TN view (q = 1) and the NN view (q = 2), float32
    and float64, each after a cache flush."""
    import torch
    dev = torch.device("cuda:0")
    flush = Flusher(torch, dev)
    gen = torch.Generator(device=dev)
    gen.manual_seed(0x77140008)
    rows = []
    for dtype in ("f32", "f64"):
        for q in (1, 2):
            for m in (32, 48, 64):
                rows.append(_timed_product(torch, args, list(EXTENTS), [1, 2, 3], q, m, False, dtype, gen, flush))
                torch.cuda.empty_cache()
    return rows


def split_and_gemv(args):
    """Products whose blocks alone would not fill the chip: the p = 1 GEMV
    (n_q = 2^24, m in {1, 16}), A [2^20, 4] with q = 1 at m = 1024, and
    A [16, 16, 2^20] with q = 3 at m = 1024 -- the skinny and split kernels."""
    import torch
    dev = torch.device("cuda:0")
    flush = Flusher(torch, dev)
    gen = torch.Generator(device=dev)
    gen.manual_seed(0x77140006)
    rows = []
    for dtype in ("f32", "f64"):
        for m in (1, 16):
            rows.append(_timed_product(torch, args, [1 << 24], [1], 1, m, False, dtype, gen, flush))
        rows.append(_timed_product(torch, args, [1 << 20, 4], [1, 2], 1, 1024, True, dtype, gen, flush))
        rows.append(_timed_product(torch, args, [16, 16, 1 << 20], [1, 2, 3], 3, 1024, True, dtype, gen, flush))
    torch.cuda.empty_cache()
    return rows


def odd_extents(args):
    """Extents that are not multiples of the vector width beside aligned
    neighbours of the same size class (float32, B column-major)."""
    import torch
    dev = torch.device("cuda:0")
    flush = Flusher(torch, dev)
    gen = torch.Generator(device=dev)
    gen.manual_seed(0x77140007)
    pairs = [(([999, 1000, 500], [1, 2, 3], 2, 16), ([1000, 1000, 500], [1, 2, 3], 2, 16)),
             (([999, 1000, 500], [1, 2, 3], 2, 1024), ([1000, 1000, 500], [1, 2, 3], 2, 1024)),
             (([1023, 1024, 512], [1, 2, 3], 1, 16), ([1024, 1024, 512], [1, 2, 3], 1, 16)),
             (([1023, 1024, 512], [1, 2, 3], 1, 1024), ([1024, 1024, 512], [1, 2, 3], 1, 1024)),
             (([1024, 1024, 513], [3, 2, 1], 3, 16), ([1024, 1024, 512], [3, 2, 1], 3, 16))]
    rows = []
    for odd, even in pairs:
        r_odd = _timed_product(torch, args, *odd, True, "f32", gen, flush)
        r_even = _timed_product(torch, args, *even, True, "f32", gen, flush)
        rows.append({"odd": r_odd, "aligned": r_even, "frac_ratio": round(r_odd["frac"] / r_even["frac"], 3)})
        torch.cuda.empty_cache()
    return rows


def host_path():
    """The drop-in's synchronous host-pointer entry on pageable NumPy buffers
    (the call stages A and B to the device and C back): one float32 product,
    A 1024x1024x512 first-order, q = 2, m = 16, timed with the host clock
    after one warm-up call.  Reported beside `value`, never as it."""
    import ttm_amd
    na, pia, q, m = list(EXTENTS), [1, 2, 3], 2, 16
    rng = np.random.default_rng(5)
    a = rng.random(int(np.prod(na)), dtype=np.float32)
    b = rng.random(m * na[q - 1], dtype=np.float32)
    nc = list(na)
    nc[q - 1] = m
    c = np.empty(int(np.prod(nc)), dtype=np.float32)
    wa, wc = ttm_amd.dense_strides(na, pia), ttm_amd.dense_strides(nc, pia)
    call = lambda: ttm_amd.ttm_raw(q, 3, a, na, wa, pia, b, [m, na[q - 1]], [2, 1], c, nc, wc)
    call()
    t0 = time.perf_counter()
    call()
    dt = time.perf_counter() - t0
    _, bytes_ = work(na, q, m, 4)
    return {"case": "f32-first-q2-m16, host (pageable) buffers", "ms": round(dt * 1e3, 2),
            "gbps_algorithmic": round(bytes_ / dt / 1e9, 1),
            "note": "includes host-to-device staging of A and B and the copy of C back over PCIe"}


def cpu_baseline(args, budget_s):
    """The oracle's `combined` restatement over OpenMP + a CBLAS, on the
    same 24-case sweep with every extent divided by 4 (A 256 x 256 x 128)."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as orc
    if not orc.LIB.exists():
        orc.build(with_ref=False)
    blas = orc.bind_blas()
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:  # pragma: no cover
        avail = os.cpu_count() or 1
    threads = max(1, min(orc.cores(), avail, args.cpu_threads))
    ext = [256, 256, 128]
    rng = np.random.default_rng(1)
    data = {}
    for dt, npdt in (("f32", np.float32), ("f64", np.float64)):
        data[dt] = rng.uniform(-1, 1, size=int(np.prod(ext))).astype(npdt)
    cases = case_list(["f32", "f64"])
    prepared = []
    for c in cases:
        nq = ext[c["q"] - 1]
        b = rng.uniform(-1, 1, size=c["m"] * nq).astype(data[c["dtype"]].dtype)
        nc = list(ext)
        nc[c["q"] - 1] = c["m"]
        prepared.append((c, b, nc, orc_strides(ext, c["pia"]), orc_strides(nc, c["pia"])))
    flops_sweep = sum(2.0 * c["m"] * int(np.prod(ext)) for c, *_ in prepared)

    case_s = [0.0] * len(prepared)

    def sweep():
        branches = []
        for i, (c, b, nc, wa, wc) in enumerate(prepared):
            t0 = time.perf_counter()
            _, br = orc.ttm_reference(c["q"], data[c["dtype"]], ext, wa, c["pia"], b, [c["m"], ext[c["q"] - 1]],
                                      [2, 1], nc, wc, use_blas=True, cores=threads)
            case_s[i] += time.perf_counter() - t0
            branches.append(br)
        return branches

    branches = sweep()  # warm-up
    case_s[:] = [0.0] * len(prepared)
    # each timed sweep follows an untimed 32 MiB scrub of the caches, as the
    # reference's measurement loop does (example/measure.cpp:85-100)
    scrub = np.empty(8 << 20, dtype=np.float32)
    times = []
    t_start = time.perf_counter()
    while True:
        scrub.fill(float(len(times)))
        t0 = time.perf_counter()
        sweep()
        times.append(time.perf_counter() - t0)
        if time.perf_counter() - t_start >= budget_s:
            break
    dt = float(np.mean(times))
    per_case = []
    for (c, *_), s in zip(prepared, case_s):
        ms = s / len(times) * 1e3
        per_case.append({"label": c["label"], "ms": round(ms, 4),
                         "tflops": round(2.0 * c["m"] * int(np.prod(ext)) / (ms * 1e-3) / 1e12, 4)})
    # the 64x48x32 products at their own size (q = 2, m = 40, both B layouts):
    # microseconds per call through the oracle's C entry (ctypes call included)
    config1 = []
    n1, m1, q1 = [64, 48, 32], 40, 2
    for dtn, npdt in (("f32", np.float32), ("f64", np.float64)):
        a1 = rng.uniform(-1, 1, size=int(np.prod(n1))).astype(npdt)
        bm = rng.uniform(-1, 1, size=(m1, n1[q1 - 1])).astype(npdt)
        nc1 = [64, m1, 32]
        for lay, pib, bflat in (("row", [2, 1], bm.reshape(-1).copy()), ("col", [1, 2], bm.T.reshape(-1).copy())):
            row = {"dtype": dtn, "b": lay}
            for key, nth in (("us_per_call", threads), ("us_per_call_1_thread", 1)):
                def call():
                    orc.ttm_reference(q1, a1, n1, [1, 64, 64 * 48], [1, 2, 3], bflat, [m1, n1[q1 - 1]], pib,
                                      nc1, [1, 64, 64 * m1], use_blas=True, cores=nth)
                for _ in range(20):
                    call()
                reps = 400
                t0 = time.perf_counter()
                for _ in range(reps):
                    call()
                row[key] = round((time.perf_counter() - t0) / reps * 1e6, 2)
            config1.append(row)
    return {
        "value": flops_sweep / dt / 1e12, "unit": "TFLOP/s", "cores": threads, "kind": "port",
        "sample": (f"the same 24-product sweep (f32+f64, first/last-order, q=1..3, m in {{16,1024}}) on A "
                   f"256x256x128 (1/64 of the volume), oracle restatement of tlib::ttm `combined` "
                   f"(OpenMP x {threads} threads + {os.path.basename(blas)}), mean of {len(times)} sweeps of "
                   f"{dt:.2f} s, each after a 32 MiB cache scrub"),
        "median_value": flops_sweep / float(np.median(times)) / 1e12,
        "blas": os.path.basename(blas), "branches": sorted(set(branches)),
        "host_cores_reference_count": orc.cores(), "hardware_concurrency": os.cpu_count(),
        "cpus_available": avail, "cases": per_case, "config1_latency": config1,
    }


def orc_strides(n, pi):
    w = [0] * len(n)
    s = 1
    for mode in pi:
        w[mode - 1] = s
        s *= n[mode - 1]
    return w


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration()
        return
    if args.force:
        import ttm_amd
        ttm_amd.force_kernel(args.force)
    cases = case_list([d.strip() for d in args.dtypes.split(",") if d.strip()])
    if args.only:
        keep = set(args.only.split(","))
        cases = [c for c in cases if c["label"] in keep]
    step_s, split, colb = gpu_bench(args, cases)
    total_flops = sum(c["flops"] for c in cases)
    total_bytes = sum(c["bytes"] for c in cases)
    kern_time = sum(c["ms"] for c in cases) * 1e-3

    # dominant kernel symbol by total time
    groups = {}
    for c in cases:
        groups.setdefault(c["symbol"], []).append(c)
    sym, members = max(groups.items(), key=lambda kv: sum(c["ms"] for c in kv[1]))
    t_sym = sum(c["ms"] for c in members) * 1e-3
    bounds = {c["bound"] for c in members}
    bound = "mfma" if bounds == {"mfma"} else "hbm"
    dt = members[0]["dtype"]
    if bound == "mfma":
        achieved = sum(c["flops"] for c in members) / t_sym / 1e12
        peak, unit = PEAK_FLOPS[dt] / 1e12, "TFLOP/s"
    else:
        achieved = sum(c["bytes"] for c in members) / t_sym / 1e9
        peak, unit = PEAK_HBM / 1e9, "GB/s"
    measured = [PMC_TRAFFIC_KIB.get(sym, {}).get(c["label"]) for c in members]
    measured = [(c, t) for c, t in zip(members, measured) if t is not None]
    traffic = traffic_alg = None
    if measured:
        traffic = float(np.mean([1024.0 * (PMC_FETCH_SCALE * f + PMC_WRITE_SCALE * wr) for _, (f, wr) in measured]))
        traffic_alg = float(np.mean([c["bytes"] for c, _ in measured]))
    roofline = {"bound": bound, "achieved": round(achieved, 3), "peak": peak, "unit": unit,
                "frac": round(achieved / peak, 4), "traffic": None if traffic is None else round(traffic),
                "kernel": sym, "launches_per_step": len(members), "avg_launch_ms": round(t_sym * 1e3 / len(members), 4),
                "cases": [c["label"] for c in members]}
    if traffic is not None:
        roofline["traffic_unit"] = "bytes per launch (HBM, rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-scaled)"
        roofline["traffic_cases"] = [c["label"] for c, _ in measured]
        roofline["traffic_over_algorithmic"] = round(traffic / traffic_alg, 3)

    tucker = tucker_batch(args) if args.tucker_count > 0 else None
    host = host_path() if args.host_path else None
    extras = None
    if args.extras:
        extras = {"config1_latency": config1_latency(args), "config4_order5": config4_order5(args),
                  "mid_m": mid_m_products(args),
                  "colmajor_b_m16": colb, "split_and_gemv": split_and_gemv(args), "odd_extents": odd_extents(args)}

    cpu = None
    if args.cpu_seconds > 0:
        cpu = cpu_baseline(args, args.cpu_seconds)

    line = {
        "metric": METRIC,
        "value": round(total_flops / kern_time / 1e12, 3),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(step_s * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "+".join(sorted({c["dtype"] for c in cases})),
        "data": "synthetic (uniform[-1,1], fixed seed)",
        "config": {"workload": "TTM sweep: A 1024x1024x512, layouts first/last-order, q in {1,2,3}, "
                               "m in {16,1024}, B m x n_q row-major",
                   "extents": EXTENTS, "layouts": list(LAYOUTS), "q": [1, 2, 3], "m": MS,
                   "products_per_step": len(cases)},
        "hbm_gbps": round(total_bytes / kern_time / 1e9, 1),
        "step_kernel_ms": round(kern_time * 1e3, 3),
        "roofline": roofline,
        "cpu_baseline": cpu,
        "bf16x3_vs_fp32": split,
        "tucker_batch": tucker,
        "host_path": host,
        "extras": extras,
        "cases": [{k: (round(v, 4) if isinstance(v, float) else v) for k, v in c.items()
                   if k in ("label", "symbol", "ms", "ms_min", "tflops", "tflops_reference_count", "gbps", "bound",
                           "frac")}
                  for c in cases],
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
