"""Batched multi-TTM chains on the GPU against the float64 definition.

The reference of every chain is the same product computed one mode at a
time with numpy.einsum in float64 (the textbook definition; the chain's
mode order does not change the exact result).  Bars: float64 1e-12 relative,
float32 normwise 4 (sum n_q) eps_32 relative to ||A|| prod ||U_k||, BF16X3
normwise-relative 1e-6.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def einsum_chain(A, Us):
    """A (natural mode order, float64) times U_k in every mode k with U_k not None."""
    X = A.astype(np.float64)
    p = X.ndim
    letters = "abcdefgh"
    for k, U in enumerate(Us):
        if U is None:
            continue
        sub = letters[:p]
        out = sub[:k] + "z" + sub[k + 1:]
        X = np.einsum(f"{sub},z{sub[k]}->{out}", X, U.astype(np.float64))
    return X


def to_layout(torch, dev, X, pi, dtype):
    """Device tensor with the values of X (natural order) stored dense in layout pi."""
    order = [m - 1 for m in reversed(pi)]               # slowest mode first
    T = torch.from_numpy(np.ascontiguousarray(np.transpose(X, order))).to(dev, dtype)
    inv = np.argsort(order)
    return T.permute(*[int(i) for i in inv])


def random_items(torch, dev, rng, count, dtype, max_ext=12, max_rank=10, orders=(1, 2, 3, 4)):
    items, refs = [], []
    for t in range(count):
        p = int(rng.choice(orders))
        na = [int(x) for x in rng.integers(1, max_ext + 1, size=p)]
        pi = [int(x) + 1 for x in rng.permutation(p)]
        A = rng.uniform(-1, 1, size=na)
        Us = []
        for k in range(p):
            if rng.random() < 0.25:
                Us.append(None)
            else:
                Us.append(rng.uniform(-1, 1, size=(int(rng.integers(1, max_rank + 1)), na[k])))
        nc = [U.shape[0] if U is not None else n for U, n in zip(Us, na)]
        At = to_layout(torch, dev, A, pi, dtype)
        Ut = [torch.from_numpy(U).to(dev, dtype).contiguous() if U is not None else None for U in Us]
        Ct = to_layout(torch, dev, np.zeros(nc), pi, dtype)
        items.append((At, Ut, Ct, pi))
        refs.append((A, Us, einsum_chain(A, Us)))
    return items, refs


def check(items, refs, dtype, label, prec=None):
    import torch
    torch.cuda.synchronize()
    eps = np.finfo(np.float32).eps
    for item, (A, Us, ref) in zip(items, refs):
        got = item[2].double().cpu().numpy()
        err = np.linalg.norm(got - ref)
        if dtype == torch.float64:
            assert err <= 1e-12 * max(np.linalg.norm(ref), 1.0), (label, A.shape)
        elif prec in ("bf16x3", "fp16x2"):
            assert err <= 1e-6 * max(np.linalg.norm(ref), 1e-30) + 1e-12, (label, A.shape, err)
        else:
            nqs = sum(n for n, U in zip(A.shape, Us) if U is not None) + sum(A.shape)
            scale = np.linalg.norm(A) * np.prod([np.linalg.norm(U) for U in Us if U is not None])
            assert err <= 4 * nqs * eps * scale, (label, A.shape, err / scale)


@pytest.mark.parametrize("dt", ["f64", "f32", "bf16x3", "fp16x2"])
def test_random_batch(gpu, dt):
    import torch
    from ttm_amd.chain import ChainPlan
    dtype = torch.float64 if dt == "f64" else torch.float32
    rng = np.random.default_rng({"f64": 1, "f32": 2, "bf16x3": 3, "fp16x2": 4}[dt])
    items, refs = random_items(torch, gpu, rng, 24, dtype)
    plan = ChainPlan(items, precision=dt if dt in ("bf16x3", "fp16x2") else None)
    _say(dt, "stage kernels", plan.stage_kernels)
    plan.run()
    check(items, refs, dtype, f"{dt} run", prec=dt)
    plan.close()


def test_modes_without_factor_are_copies(gpu):
    """A mode whose factor is None keeps its extent through a copy stage
    (k_transpose_grouped), exactly: a long unfactored mode (60000, whose
    identity would take 14 GB), every mode unfactored (the chain is a
    layout round trip), and unfactored modes between factored ones, in both
    element types; the copy stages are bit-exact, so a chain of copies
    returns A unchanged."""
    import torch
    from ttm_amd.chain import ChainPlan
    rng = np.random.default_rng(21)
    for dtype in (torch.float32, torch.float64):
        specs = [([4, 3, 60000], [2, 1, 3], [True, True, False]),
                 ([7, 130, 5, 9], [3, 1, 4, 2], [False, False, False, False]),
                 ([33, 65, 17], [1, 2, 3], [False, True, False]),
                 ([70, 3, 66], [3, 2, 1], [True, False, True])]
        items, refs = [], []
        for na, pi, has in specs:
            A = rng.uniform(-1, 1, size=na)
            Us = [rng.uniform(-1, 1, size=(int(rng.integers(1, 9)), n)) if h else None for n, h in zip(na, has)]
            nc = [U.shape[0] if U is not None else n for U, n in zip(Us, na)]
            At = to_layout(torch, gpu, A, pi, dtype)
            Ut = [torch.from_numpy(U).to(gpu, dtype).contiguous() if U is not None else None for U in Us]
            Ct = to_layout(torch, gpu, np.full(nc, np.nan), pi, dtype)
            items.append((At, Ut, Ct, pi))
            refs.append((A, Us, einsum_chain(A, Us)))
        plan = ChainPlan(items)
        kernels = plan.stage_kernels
        _say(dtype, "stage kernels", kernels)
        assert any(k.startswith("k_transpose_grouped<") for k in kernels)
        plan.run()
        torch.cuda.synchronize()
        # all modes unfactored: C is A, bit for bit, in the same layout
        assert torch.equal(items[1][2], items[1][0])
        check(items, refs, dtype, f"copy stages {dtype}")
        plan.close()


def _short_row_items(torch, dev, rng, dtype):
    """Chains whose stages the short-row kernel can take (contractions
    4..256, ranks 1..64): orders 2 to 5, every layout, contraction lengths
    at 4, 5, 15..17, 255, 256 and random, rank classes 16 / 32 / 48 / 64 at their
    edges, single columns, and 200 small tensors (more problems than
    workgroups, so workgroup ranges cross many problem boundaries)."""
    specs = [([4, 256], 1), ([256, 5], 64), ([17, 15, 16], 33), ([255, 7, 9], 16), ([64, 65, 4], 17),
             ([16, 31, 33], 32), ([200, 130], 48), ([4, 4, 4], 4)]
    for _ in range(200):
        p = int(rng.integers(2, 4))
        specs.append(([int(x) for x in rng.integers(4, 41 if p == 3 else 257, size=p)], int(rng.integers(1, 65))))
    for _ in range(16):  # orders 4 and 5: longer chains, every stage on the short-row kernels
        p = int(rng.integers(4, 6))
        specs.append(([int(x) for x in rng.integers(4, 13, size=p)], int(rng.integers(1, 17))))
    items, refs = [], []
    for na, rmax in specs:
        p = len(na)
        pi = [int(x) + 1 for x in rng.permutation(p)]
        A = rng.uniform(-1, 1, size=na)
        Us = [rng.uniform(-1, 1, size=(int(rng.integers(1, rmax + 1)) if k else rmax, n)) for k, n in enumerate(na)]
        nc = [U.shape[0] for U in Us]
        At = to_layout(torch, dev, A, pi, dtype)
        Ut = [torch.from_numpy(U).to(dev, dtype).contiguous() for U in Us]
        Ct = to_layout(torch, dev, np.full(nc, np.nan), pi, dtype)
        items.append((At, Ut, Ct, pi))
        refs.append((A, Us, einsum_chain(A, Us)))
    return items, refs


@pytest.mark.parametrize("kind", ["f32", "bf16x3", "fp16x2", "f64"])
def test_short_row_kernel_matches_the_definition(gpu, monkeypatch, kind):
    """Stages with contractions of 4..256 and ranks up to 64 run the
    short-row kernels (k_rows_grouped, k_rows_bf16x3, k_rows_fp16x2,
    k_rows_f64) by default in float32, BF16X3, FP16X2 and float64, one
    launch per stage for a batch this small; with TTM_AMD_CHAIN_MERGE=0 one
    launch per rank class (every row height); with TTM_AMD_CHAIN_ROWS=0 the
    same plan runs the grouped tiles.  Both are
    checked against the float64 definition (every output written: C starts
    as NaN), eagerly and as a graph replay, and against each other.  The
    split precisions are held to the float32 bar here (normwise error
    relative to ||A|| prod ||U_k||): the batch has cores of one element that
    cancel hundreds of terms, where an error relative to the result itself
    measures the cancellation, not the arithmetic."""
    import torch
    from ttm_amd.chain import ChainPlan
    rng = np.random.default_rng(31)
    dtype = torch.float64 if kind == "f64" else torch.float32
    prec = kind if kind in ("bf16x3", "fp16x2") else None
    items, refs = _short_row_items(torch, gpu, rng, dtype)
    name = {"f32": "k_rows_grouped<", "bf16x3": "k_rows_bf16x3<", "fp16x2": "k_rows_fp16x2<", "f64": "k_rows_f64<"}[kind]
    # by default every stage of this small batch is one launch on the
    # tallest row height it needs
    plan = ChainPlan(items, precision=prec)
    _say("merged short-row stage kernels", plan.stage_kernels)
    stages = max(len(it[0].shape) for it in items)
    assert len(plan.stage_kernels) == stages and all(k.startswith(name) for k in plan.stage_kernels)
    plan.run()
    check(items, refs, dtype, "short rows, merged classes")
    plan.close()
    # every rank class in a launch of its own: each row height on the edge cases
    monkeypatch.setenv("TTM_AMD_CHAIN_MERGE", "0")
    for it in items:
        it[2].fill_(float("nan"))
    plan = ChainPlan(items, precision=prec)
    kernels = plan.stage_kernels
    _say("short-row stage kernels", kernels)
    for h in (16, 32, 48, 64):
        assert any(k.startswith(f"{name}{h},") for k in kernels), h
    plan.run()
    check(items, refs, dtype, "short rows")
    got_rows = [it[2].clone() for it in items]
    for it in items:
        it[2].fill_(float("nan"))
    plan.run_graph()
    check(items, refs, dtype, "short rows graph")
    plan.close()

    monkeypatch.setenv("TTM_AMD_CHAIN_ROWS", "0")
    for it in items:
        it[2].fill_(float("nan"))
    plan = ChainPlan(items, precision=prec)
    assert not any(k.startswith("k_rows_") for k in plan.stage_kernels)
    plan.run()
    check(items, refs, dtype, "grouped tiles")
    eps = np.finfo(np.float64 if kind == "f64" else np.float32).eps
    for a, it, (A, Us, _) in zip(got_rows, items, refs):
        scale = np.linalg.norm(A) * np.prod([np.linalg.norm(U) for U in Us])
        diff = float(torch.linalg.vector_norm((a - it[2]).double()))
        assert diff <= 8 * sum(A.shape) * eps * scale, (A.shape, diff / scale)
    plan.close()


@pytest.mark.parametrize("kind", ["f32", "bf16x3", "fp16x2", "f64"])
def test_short_row_48_and_64_row_launches(gpu, kind):
    """A first stage big enough (36 tensors 40 x 128 x 128: 128 units of 16
    columns per compute unit and more) that ranks 33..48 and 49..64 take two
    launches, the 48-row and the 64-row short-row kernel; the later, smaller
    stages are one launch each.  Against the float64 definition at the
    float32 (float64) bar, eagerly and as a graph replay."""
    import torch
    from ttm_amd.chain import ChainPlan
    rng = np.random.default_rng(53)
    dtype = torch.float64 if kind == "f64" else torch.float32
    prec = kind if kind in ("bf16x3", "fp16x2") else None
    items, refs = [], []
    for t in range(36):
        na = [40, 128, 128]
        r1 = int(rng.integers(33, 49)) if t % 2 == 0 else int(rng.integers(49, 65))
        ranks = [r1, int(rng.integers(4, 17)), int(rng.integers(4, 17))]
        A = rng.uniform(-1, 1, size=na)
        Us = [rng.uniform(-1, 1, size=(r, n)) for r, n in zip(ranks, na)]
        pi = [1, 2, 3]
        At = to_layout(torch, gpu, A, pi, dtype)
        Ut = [torch.from_numpy(U).to(gpu, dtype).contiguous() for U in Us]
        Ct = to_layout(torch, gpu, np.full(ranks, np.nan), pi, dtype)
        items.append((At, Ut, Ct, pi))
        refs.append((A, Us, einsum_chain(A, Us)))
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    assert 36 * (128 * 128 // 16) >= 128 * cus
    name = {"f32": "k_rows_grouped<", "bf16x3": "k_rows_bf16x3<", "fp16x2": "k_rows_fp16x2<", "f64": "k_rows_f64<"}[kind]
    plan = ChainPlan(items, precision=prec)
    kernels = plan.stage_kernels
    _say(kind, "48/64-row stage kernels", kernels)
    assert any(k.startswith(name + "48,") for k in kernels)
    assert any(k.startswith(name + "64,") for k in kernels)
    plan.run()
    check(items, refs, dtype, f"48/64-row launches {kind}")
    for it in items:
        it[2].fill_(float("nan"))
    plan.run_graph()
    check(items, refs, dtype, f"48/64-row launches {kind} graph")
    plan.close()


@pytest.mark.parametrize("prec", ["bf16x3", "fp16x2"])
def test_short_row_split_precisions_far_from_one(gpu, prec):
    """The split short-row kernels on tensors and factors scaled by powers of
    two far outside the binary16 range (A by 2^-40 .. 2^40, the factors by
    2^-20 .. 2^20; every result stays a normal float32): FP16X2 scales A per
    unit and B per problem before the split, BF16X3 needs no scale; both stay
    at the 1e-6 normwise bar against the float64 definition."""
    import torch
    from ttm_amd.chain import ChainPlan
    rng = np.random.default_rng(47)
    items, refs = [], []
    for t in range(24):
        na = [int(x) for x in rng.integers(4, 70, size=3)]
        pi = [int(x) + 1 for x in rng.permutation(3)]
        sa, su = 2.0 ** int(rng.integers(-40, 41)), 2.0 ** int(rng.integers(-20, 21))
        A = rng.uniform(-1, 1, size=na) * sa
        Us = [rng.uniform(-1, 1, size=(int(rng.integers(1, 65)), n)) * su for n in na]
        At = to_layout(torch, gpu, A, pi, torch.float32)
        Ut = [torch.from_numpy(U).to(gpu, torch.float32).contiguous() for U in Us]
        Ct = to_layout(torch, gpu, np.full([U.shape[0] for U in Us], np.nan), pi, torch.float32)
        # the reference takes the float32-rounded inputs (natural mode order)
        A32 = At.double().cpu().numpy()
        items.append((At, Ut, Ct, pi))
        refs.append((A32, [U.double().cpu().numpy() for U in Ut], einsum_chain(A32, [U.double().cpu().numpy() for U in Ut])))
    plan = ChainPlan(items, precision=prec)
    kernels = plan.stage_kernels
    _say(prec, "far-from-one stage kernels", kernels)
    assert all(k.startswith("k_rows_") for k in kernels)
    plan.run()
    check(items, refs, torch.float32, f"{prec} far from one", prec=prec)
    plan.close()


@pytest.mark.parametrize("kind", ["f32", "bf16x3", "fp16x2", "f64"])
def test_mixed_kernel_batch(gpu, kind):
    """One plan whose stages mix the short-row kernels and the tiles: rank
    classes above 64 and contractions above 256 or below 4 (tiles) next to
    eligible ones, orders 1 to 4, modes without a factor (copy stages), every
    layout; against the float64 definition at the float32 (float64) bar."""
    import torch
    from ttm_amd.chain import ChainPlan
    rng = np.random.default_rng(53)
    dtype = torch.float64 if kind == "f64" else torch.float32
    prec = kind if kind in ("bf16x3", "fp16x2") else None
    items, refs = [], []
    for t in range(48):
        p = int(rng.integers(1, 5))
        hi = {1: 600, 2: 300, 3: 40, 4: 12}[p]
        na = [int(x) for x in rng.integers(1, hi + 1, size=p)]
        pi = [int(x) + 1 for x in rng.permutation(p)]
        A = rng.uniform(-1, 1, size=na)
        Us = []
        for k in range(p):
            if rng.random() < 0.15:
                Us.append(None)
            else:
                r = int(rng.integers(1, 101)) if p <= 2 else int(rng.integers(1, 17))
                Us.append(rng.uniform(-1, 1, size=(r, na[k])))
        nc = [U.shape[0] if U is not None else n for U, n in zip(Us, na)]
        At = to_layout(torch, gpu, A, pi, dtype)
        Ut = [torch.from_numpy(U).to(gpu, dtype).contiguous() if U is not None else None for U in Us]
        Ct = to_layout(torch, gpu, np.full(nc, np.nan), pi, dtype)
        items.append((At, Ut, Ct, pi))
        refs.append((A, Us, einsum_chain(A, Us)))
    plan = ChainPlan(items, precision=prec)
    kernels = plan.stage_kernels
    _say(kind, "mixed stage kernels", kernels)
    assert any(k.startswith("k_rows_") for k in kernels)
    assert any(not k.startswith("k_rows_") and not k.startswith("k_transpose") for k in kernels)
    plan.run_graph()
    check(items, refs, dtype, f"mixed {kind}")
    plan.close()


def test_graph_replay_matches_eager(gpu):
    """Replays of the captured chain, with eager work in between, equal the eager run."""
    import torch
    from ttm_amd.chain import ChainPlan
    rng = np.random.default_rng(7)
    items, refs = random_items(torch, gpu, rng, 16, torch.float32, max_ext=20, max_rank=12, orders=(3,))
    plan = ChainPlan(items)
    plan.run()
    torch.cuda.synchronize()
    eager = [it[2].clone() for it in items]
    for rep in range(4):
        for it in items:
            it[2].fill_(float("nan"))
        torch.cuda.synchronize()
        plan.run_graph()
        torch.cuda.synchronize()
        junk = torch.randn(1000, device=gpu).sum()  # eager work between replays
        for it, E in zip(items, eager):
            assert torch.equal(it[2], E), f"replay {rep}"
        del junk
    check(items, refs, torch.float32, "graph")
    plan.close()


def test_tune_keeps_results(gpu, tmp_path, monkeypatch):
    import torch
    from ttm_amd.chain import ChainPlan
    monkeypatch.setenv("TTM_AMD_TUNE_CACHE", str(tmp_path / "tune.txt"))
    rng = np.random.default_rng(9)
    items, refs = random_items(torch, gpu, rng, 12, torch.float32, max_ext=40, max_rank=40, orders=(3,))
    plan = ChainPlan(items)
    plan.tune()
    _say("tuned stage kernels", plan.stage_kernels)
    for it in items:
        it[2].zero_()
    plan.run()
    check(items, refs, torch.float32, "tuned")
    plan.close()


def orthonormal_rows(rng, r, n):
    Q, _ = np.linalg.qr(rng.standard_normal((n, r)))
    return Q[:, :r].T.copy()  # r x n with orthonormal rows


@pytest.mark.parametrize("prec", [None, "bf16x3", "fp16x2"])
def test_tucker_batch_reduced(gpu, prec):
    """Tucker chains A x1 U1 x2 U2 x3 U3 over 32 uneven first-order tensors (extents 16..64, ranks 4..16)."""
    import torch
    from ttm_amd.chain import ChainPlan
    rng = np.random.default_rng(0x77140005)
    items, refs = [], []
    for t in range(32):
        na = [int(x) for x in rng.integers(16, 65, size=3)]
        rk = [int(x) for x in rng.integers(4, 17, size=3)]
        A = rng.uniform(-1, 1, size=na)
        Us = [orthonormal_rows(rng, r, n) for r, n in zip(rk, na)]
        At = to_layout(torch, gpu, A, [1, 2, 3], torch.float32)
        Ut = [torch.from_numpy(U).to(gpu, torch.float32).contiguous() for U in Us]
        Ct = to_layout(torch, gpu, np.zeros(rk), [1, 2, 3], torch.float32)
        items.append((At, Ut, Ct))
        refs.append((A, Us, einsum_chain(A.astype(np.float32), [U.astype(np.float32) for U in Us])))
    plan = ChainPlan(items, precision=prec)
    plan.run_graph()
    check(items, refs, torch.float32, f"tucker {prec}", prec=prec)
    plan.close()


def _tensordot_chain(A, Us):
    """A (natural mode order) x_k U_k for k = 1, 2, 3 in float64 through BLAS."""
    X = A.astype(np.float64)
    for k, U in enumerate(Us):
        X = np.moveaxis(np.tensordot(U.astype(np.float64), X, axes=([1], [k])), 0, k)
    return X


@pytest.mark.parametrize("prec", [None, "bf16x3", "fp16x2", "f64"])
def test_tucker_batch_config5_full_size(gpu, prec):
    """The Tucker batch of the benchmark at its stated size: 256 first-order
    float32 tensors, extents U{16..256}^3, ranks U{4..64}, orthonormal-row
    factors (bench.tucker_items, the same seed), as one hipGraph replay of a
    plan (tuned for FP32), every core against a float64 tensordot chain on
    the host; "f64" runs the same tensors and factors converted to float64
    (the bench's FP64 leg), tuned, at the float64 bar."""
    import torch
    import bench
    from ttm_amd.chain import ChainPlan
    items, ranks, flops, _ = bench.tucker_items(torch, gpu, 256)
    if prec == "f64":
        items = [(A.double(), [U.double() for U in Us], C.double()) for A, Us, C in items]
    _say(f"config 5 {prec or 'fp32'}: {len(items)} tensors, {flops / 1e9:.1f} GFLOP")
    plan = ChainPlan(items, precision=prec if prec != "f64" else None)
    if prec in (None, "f64"):
        plan.tune()
    for it in items:
        it[2].fill_(float("nan"))
    plan.run_graph()
    torch.cuda.synchronize()
    _say("stage kernels", plan.stage_kernels)
    eps = np.finfo(np.float64 if prec == "f64" else np.float32).eps
    worst = 0.0
    for t, (A, Us, C) in enumerate(items):
        a = A.double().cpu().numpy()
        us = [U.double().cpu().numpy() for U in Us]
        ref = _tensordot_chain(a, us)
        got = C.double().cpu().numpy()
        err = float(np.linalg.norm(got - ref))
        if prec in (None, "f64"):
            nqs = sum(a.shape) + sum(a.shape)
            scale = np.linalg.norm(a) * np.prod([np.linalg.norm(U) for U in us])
            bound = 4 * nqs * eps * scale
        else:
            bound = 1e-6 * max(np.linalg.norm(ref), 1e-30)
        worst = max(worst, err / bound)
        assert np.isfinite(got).all() and err <= bound, (t, a.shape, ranks[t], err, bound)
    _say(f"config 5 {prec or 'fp32'}: worst error / bound {worst:.3e}")
    plan.close()
    del items
    torch.cuda.empty_cache()
