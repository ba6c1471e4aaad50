"""Planner properties on random metadata (host logic only, no GPU work).

For every valid problem the planner must name a family whose preconditions the
problem meets: the streaming kernels hold all m rows of C in one wave (m <= 64)
and read whole 16-byte vectors along the contiguous block (NN: w % vec == 0;
TN: w == 1, n_q % vec == 0); `k_tn_lds` keeps all of B in LDS, so its B stride
must cover ceil(n_q / vec) pieces rounded to 32, plus one, and B plus the waves'
8 KB buffers must fit the 160 KB of LDS of a gfx950 CU.  A forced family whose
preconditions a problem does not meet is ignored for it.  Invalid metadata
(the reference's validation failures, include/tlib/ttm.h:65-90) plans nothing.
The one-thread-per-element kernel runs only when forced or for an A of fewer
elements than one 16-byte vector: every other valid dense problem gets a
streaming, skinny or tile kernel, the NN tile padding its slabs exactly when
w % vec != 0.
"""
import re

import numpy as np
import pytest

FAMILIES = {"generic", "nn_stream", "tn_stream", "gemm", "lowp", "lowp_generic", "skinny", "none"}
LDS_BYTES = 160 * 1024


def _fold(q, na, pia):
    qh = pia.index(q) + 1
    w = int(np.prod([na[r - 1] for r in pia[:qh - 1]], dtype=np.int64)) if qh > 1 else 1
    return w, na[q - 1]


def _random_problem(rng):
    p = int(rng.integers(1, 6))
    na = []
    for _ in range(p):
        r = rng.random()
        na.append(int(rng.integers(1, 9)) * (4 if r < 0.5 else 1) if r < 0.8 else int(rng.integers(100, 2100)))
    pia = [int(x) + 1 for x in rng.permutation(p)]
    q = int(rng.integers(1, p + 1))
    m = int(rng.choice([1, 2, 4, 8, 15, 16, 17, 32, 63, 64, 65, 128, 1024]))
    pib = [1, 2] if rng.random() < 0.5 else [2, 1]
    return q, na, pia, [m, na[q - 1]], pib


def _check(q, na, pia, nb, pib, dbl, prec, k, sym, forced=False):
    m = nb[0]
    vec = 2 if dbl else 4
    w, nq = _fold(q, na, pia)
    outer = int(np.prod(na, dtype=np.int64)) // (w * nq)
    case = (q, na, pia, nb, pib, dbl, prec, k, sym)
    assert k in FAMILIES, case
    if dbl and prec != 0:
        assert k == "none", case
        return
    if prec != 0:
        assert k in ("lowp", "lowp_generic"), case
        return
    assert k not in ("lowp", "lowp_generic", "none"), case
    if not forced and int(np.prod(na, dtype=np.int64)) >= vec:
        assert k != "generic", case  # only an A of fewer elements than a vector
    if k == "skinny":
        assert w == 1 and outer <= 8 and nq >= 512, case
    if k == "gemm":
        pad = sym.rstrip(">").split(", ")[-1]
        assert pad == ("true" if (w != 1 and w % vec != 0) else "false"), case
    if k == "nn_stream":
        assert m <= 64 and w % vec == 0, case
    if k == "tn_stream":
        assert m <= 64 and w == 1 and nq % vec == 0, case
    if sym.startswith("k_tn_lds<"):
        waves, depth, spb_max = (int(x) for x in re.findall(r"\d+", sym.split(",", 1)[1]))
        spb = ((nq // vec + 31) // 32) * 32 + 1
        assert m <= 16 and spb <= spb_max, case
        assert 16 * spb_max * 16 + waves * 8192 <= LDS_BYTES, case


def test_planner_preconditions_on_random_problems(ttmlib):
    import ttm_amd
    rng = np.random.default_rng(7)
    seen = set()
    for _ in range(3000):
        q, na, pia, nb, pib = _random_problem(rng)
        dbl = bool(rng.random() < 0.4)
        prec = int(rng.choice([0, 0, 0, 1, 2, 3, 4, 5]))
        k = ttm_amd.plan_kernel(q, na, pia, nb, pib, is_double=dbl, precision=prec)
        sym = ttm_amd.plan_kernel_symbol(q, na, pia, nb, pib, is_double=dbl, precision=prec)
        _check(q, na, pia, nb, pib, dbl, prec, k, sym)
        seen.add(sym.split("<")[0] if sym else k)
    # the random shapes reach every plain-precision family the planner picks
    assert {"k_nn_stream", "k_tn_lds", "k_gemm"} <= seen, seen


@pytest.mark.parametrize("family", ["generic", "nn_stream", "tn_stream", "gemm", "skinny"])
def test_forced_family_only_where_it_can_run(ttmlib, family):
    import ttm_amd
    rng = np.random.default_rng(11)
    hits = 0
    ttm_amd.force_kernel(family)
    try:
        for _ in range(1500):
            q, na, pia, nb, pib = _random_problem(rng)
            dbl = bool(rng.random() < 0.4)
            k = ttm_amd.plan_kernel(q, na, pia, nb, pib, is_double=dbl)
            sym = ttm_amd.plan_kernel_symbol(q, na, pia, nb, pib, is_double=dbl)
            _check(q, na, pia, nb, pib, dbl, 0, k, sym, forced=True)
            hits += k == family
    finally:
        ttm_amd.force_kernel(None)
    assert hits > 0, family
    if family == "generic":
        assert hits == 1500


def test_invalid_metadata_plans_nothing(ttmlib):
    import ttm_amd
    rng = np.random.default_rng(5)
    for _ in range(1000):
        q, na, pia, nb, pib = _random_problem(rng)
        p = len(na)
        kind = int(rng.integers(0, 5))
        if kind == 0 and p > 1:
            pia = list(pia)
            pia[0] = pia[int(rng.integers(1, p))]               # a repeated mode
        elif kind == 1:
            nb = [nb[0], nb[1] + int(rng.integers(1, 5))]      # nb[1] != na[q-1]
        elif kind == 2:
            q = p + int(rng.integers(1, 3))                     # q > p
        elif kind == 3:
            pia = list(pia)
            pia[int(rng.integers(0, p))] = p + 1                # mode out of range
        else:
            na = list(na)
            na[int(rng.integers(0, p))] = 0                     # empty extent
        assert ttm_amd.plan_kernel(q, na, pia, nb, pib) == "", (kind, q, na, pia, nb)
        assert ttm_amd.plan_kernel_symbol(q, na, pia, nb, pib) == "", (kind, q, na, pia, nb)
