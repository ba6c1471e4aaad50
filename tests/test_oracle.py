"""Pins the CPU oracle (oracle/ttm_oracle.c) before anything is compared with it.

* metadata: the reference test-suite's literal known answers, and the real
  reference headers compiled under oracle/_ref on exhaustive small inputs;
* contraction: the reference's closed-form sweep (all p! layouts, every q,
  both B layouts, orders 2..4) plus a B that depends on both indices,
  numpy.einsum in float64 on random data, and the ttmpy test vectors;
* a planted mistake (transposed B) must fail the closed-form check.
"""
import itertools
import json
import pathlib

import numpy as np
import pytest

import refcases as rc

GOLD = json.loads((pathlib.Path(__file__).parent / "golden" / "reference_known_answers.json").read_text())


# ----------------------------------------------------------------------------- metadata
def test_shape_truth_tables(oracle):
    t = GOLD["shape_table"]
    L = oracle.lib()
    for i, s in enumerate(t["shapes"]):
        n = oracle.sz(s)
        assert bool(L.or_is_scalar(n, len(s))) == (i in t["is_scalar"]), s
        assert bool(L.or_is_vector(n, len(s))) == (i in t["is_vector"]), s
        assert bool(L.or_is_matrix(n, len(s))) == (i in t["is_matrix"]), s
        assert bool(L.or_is_tensor(n, len(s))) == (i in t["is_tensor"]), s
        assert bool(L.or_is_valid_shape(n, len(s))) == (i not in t["invalid"]), s


def test_k_order_layouts(oracle):
    L = oracle.lib()
    for k, layouts in GOLD["k_order_layouts"].items():
        if k.startswith("_"):
            continue
        for ref in layouts:
            got = oracle.k_order_layout(len(ref), int(k))
            assert got == ref
            assert L.or_is_valid_layout(oracle.sz(got), len(got))
    # k = 0 is the last order
    for p in range(1, 11):
        assert oracle.k_order_layout(p, 0) == list(range(p, 0, -1))
        assert L.or_is_last_order(oracle.sz(list(range(p, 0, -1))), p)
        assert L.or_is_first_order(oracle.sz(list(range(1, p + 1))), p)


def test_invalid_layouts(oracle):
    L = oracle.lib()
    for lay in GOLD["invalid_layouts"]["layouts"]:
        assert not L.or_is_valid_layout(oracle.sz(lay), len(lay)), lay


def test_inverse_mode_roundtrip(oracle):
    # test/src/gtest_tlib_layout.cpp:193-211
    for order in range(1, 11):
        for fmt in range(1, order + 1):
            lay = oracle.k_order_layout(order, fmt)
            for mode in range(1, order + 1):
                r = oracle.inverse_mode(lay, mode)
                assert 1 <= r <= order and lay[r - 1] == mode
    assert oracle.inverse_mode([1, 2, 3], 0) == 0
    assert oracle.inverse_mode([1, 2, 3], 4) == 0
    assert oracle.inverse_mode([1, 1, 3], 1) == 0


def test_strides_known_answers(oracle):
    L = oracle.lib()
    for c in GOLD["strides"]["cases"]:
        w = oracle.compute_strides(c["shape"], c["layout"])
        assert w == c["strides"], c
        assert L.or_is_valid_strides(oracle.sz(c["layout"]), len(w), oracle.sz(w)) == 1


def test_case_examples(oracle):
    for e in GOLD["mtm_cases"]["examples"]:
        hits = [k for k in range(1, 9) if oracle.is_case(k, e["p"], e["q"], e["pi"])]
        assert hits == [e["case"]], e


@pytest.mark.skipif(not pathlib.Path(__file__).resolve().parents[1].joinpath(
    "oracle", "_ref", "libttm_ref_meta.so").exists() and not pathlib.Path("/root/reference/include").is_dir(),
    reason="reference sources not available to build oracle/_ref")
def test_metadata_matches_compiled_reference(oracle):
    """Exhaustive comparison with the reference's own headers (oracle/_ref)."""
    R = oracle.ref_lib()
    if R is None:
        oracle.build(with_ref=True)
        R = oracle.ref_lib()
    assert R is not None
    L = oracle.lib()
    sz = oracle.sz
    # shapes: every tuple over {0,1,2,3} up to order 4
    for p in range(0, 5):
        for s in itertools.product(range(4), repeat=p):
            n = sz(s)
            for f in ("is_valid_shape", "is_scalar", "is_vector", "is_matrix", "is_tensor"):
                assert getattr(L, "or_" + f)(n, p) == getattr(R, "ref_" + f)(n, p), (f, s)
    # layouts: every tuple over {0..p+1} up to order 4 (valid and invalid)
    for p in range(1, 5):
        for lay in itertools.product(range(p + 2), repeat=p):
            pl = sz(lay)
            assert L.or_is_valid_layout(pl, p) == R.ref_is_valid_layout(pl, p), lay
            assert L.or_is_first_order(pl, p) == R.ref_is_first_order(pl, p), lay
            assert L.or_is_last_order(pl, p) == R.ref_is_last_order(pl, p), lay
            for mode in range(0, p + 2):
                assert L.or_inverse_mode(pl, p, mode) == R.ref_inverse_mode(pl, p, mode), (lay, mode)
            for case in range(1, 9):
                for q in range(1, p + 1):
                    assert L.or_is_case(case, p, q, pl) == R.ref_is_case(case, p, q, pl), (case, lay, q)
    # k-order layouts
    for p in range(1, 8):
        for k in range(0, p + 3):
            a, b = sz([0] * p), sz([0] * p)
            L.or_compute_k_order_layout(a, p, k)
            R.ref_compute_k_order_layout(b, p, k)
            assert list(a) == list(b), (p, k)
    # strides and their validity on all shapes over {1,2,3} x all layouts, p <= 4
    rng = np.random.default_rng(7)
    for p in range(1, 5):
        for s in itertools.product((1, 2, 3), repeat=p):
            for lay in itertools.permutations(range(1, p + 1)):
                w1, w2, w3 = sz([0] * p), sz([0] * p), sz([0] * p)
                assert L.or_compute_strides(sz(s), sz(lay), p, w1) == R.ref_compute_strides(sz(s), sz(lay), p, w2)
                assert list(w1) == list(w2), (s, lay)
                assert R.ref_tensor_strides(sz(s), sz(lay), p, w3) == 0 and list(w3) == list(w2)
                for _ in range(3):
                    rw = sz(rng.integers(1, 6, size=p).tolist())
                    assert L.or_is_valid_strides(sz(lay), p, rw) == R.ref_is_valid_strides(sz(lay), p, rw)


# ----------------------------------------------------------------------------- contraction
def _run_closed_form(oracle, cases, dtype):
    n = 0
    for na, pia, q, colmajor, bk in cases:
        m = 2 * na[q - 1]  # the reference tests use B of shape 2 n_q x n_q
        a, wa = rc.ttm_init(q, na, pia)
        b, pib = rc.make_b(m, na[q - 1], bk, colmajor, dtype)
        nc = list(na)
        nc[q - 1] = m
        wc = rc.dense_strides(nc, pia)
        c, _ = oracle.ttm_reference(q, a.astype(dtype), na, wa, pia, b, [m, na[q - 1]], pib, nc, wc)
        expect = rc.closed_form(q, na, pia, m, bk)
        got = rc.as_nd(c, nc, wc)
        assert np.array_equal(got, expect), (na, pia, q, colmajor, bk)
        n += 1
    return n


def test_closed_form_sweep_f64(oracle):
    n = _run_closed_form(oracle, rc.closed_form_cases(orders=(2, 3, 4)), np.float64)
    assert n > 10000


def test_closed_form_sweep_f32(oracle):
    # B = ones keeps every sum below 2^24, so float32 is exact as well
    n = _run_closed_form(oracle, rc.closed_form_cases(orders=(2, 3), b_kinds=("ones",)), np.float32)
    assert n > 100


def test_order1_and_order5(oracle):
    cases = [([7], [1], 1, cm, bk) for cm in (False, True) for bk in ("ones", "jl")]
    cases += [(na, pia, q, cm, "jl") for na in ([2, 3, 2, 2, 3],) for pia in rc.generate_permutations(5)[::7]
              for q in range(1, 6) for cm in (False, True)]
    assert _run_closed_form(oracle, cases, np.float64) == len(cases)


@pytest.mark.parametrize("cores", [1, 4, 64, 100000])
def test_all_combined_branches_vs_einsum(oracle, cores):
    """Every branch of `combined` (the choice depends on the core count) on random data."""
    rng = np.random.default_rng(cores)
    letters = "abcdefgh"
    branches = set()
    for na, pia in [([5, 6, 7], [1, 2, 3]), ([5, 6, 7], [3, 1, 2]), ([3, 4, 5, 6], [2, 4, 1, 3]),
                    ([2, 3, 4, 5, 3], [5, 1, 3, 2, 4]), ([9], [1]), ([4, 7], [2, 1])]:
        p = len(na)
        for q in range(1, p + 1):
            for colmajor in (False, True):
                m = 3 + q
                A = rng.uniform(-1, 1, size=na)
                Bm = rng.uniform(-1, 1, size=(m, na[q - 1]))
                wa = rc.dense_strides(na, pia)
                # flat storage of A in layout pia
                a = np.zeros(A.size)
                rc.as_nd(a, na, wa)[...] = A
                b = (Bm.T if colmajor else Bm).reshape(-1).copy()
                nc = list(na)
                nc[q - 1] = m
                wc = rc.dense_strides(nc, pia)
                c, br = oracle.ttm_reference(q, a, na, wa, pia, b, [m, na[q - 1]], [1, 2] if colmajor else [2, 1],
                                             nc, wc, cores=cores)
                branches.add(br)
                sub = letters[:p]
                out = sub[: q - 1] + "z" + sub[q:]
                expect = np.einsum(f"{sub},z{sub[q - 1]}->{out}", A, Bm)
                np.testing.assert_allclose(rc.as_nd(c, nc, wc), expect, rtol=1e-12, atol=1e-12)
    assert 0 in branches


def test_brute_force_vs_einsum(oracle):
    rng = np.random.default_rng(3)
    for na, pia, q in [([4, 5, 6], [2, 3, 1], 2), ([3, 4, 2, 5], [4, 3, 2, 1], 4), ([6], [1], 1)]:
        m = 5
        A = rng.uniform(-1, 1, size=na)
        Bm = rng.uniform(-1, 1, size=(m, na[q - 1]))
        wa = rc.dense_strides(na, pia)
        a = np.zeros(A.size)
        rc.as_nd(a, na, wa)[...] = A
        nc = list(na)
        nc[q - 1] = m
        wc = rc.dense_strides(nc, pia)
        for colmajor in (False, True):
            b = (Bm.T if colmajor else Bm).reshape(-1).copy()
            c = oracle.ttm_brute_f64(q, a, na, wa, b, m, colmajor, nc, wc)
            p = len(na)
            sub = "abcdefgh"[:p]
            expect = np.einsum(f"{sub},z{sub[q - 1]}->{sub[:q - 1]}z{sub[q:]}", A, Bm)
            np.testing.assert_allclose(rc.as_nd(c, nc, wc), expect, rtol=1e-13, atol=1e-13)


def test_ttmpy_vectors_through_oracle(oracle):
    """ttmpy/tests/test.py:7-27 -- exact equality with numpy.einsum (last-order A, row-major B)."""
    for case in GOLD["ttmpy_einsum"]["ttm"]:
        A = np.arange(np.prod(case["a_shape"]), dtype=np.float64).reshape(case["a_shape"])
        B = np.arange(np.prod(case["b_shape"]), dtype=np.float64).reshape(case["b_shape"])
        q = case["q"]
        p = A.ndim
        pia = list(range(p, 0, -1))
        wa = [s // 8 for s in A.strides]
        nc = list(A.shape)
        nc[q - 1] = B.shape[0]
        wc = rc.dense_strides(nc, pia)
        c, _ = oracle.ttm_reference(q, A.reshape(-1), list(A.shape), wa, pia, B.reshape(-1), list(B.shape),
                                    [2, 1], nc, wc)
        assert np.array_equal(c.reshape(nc), np.einsum(case["einsum"], A, B))


def test_validation_order(oracle):
    """include/tlib/ttm.h:65-90: the first failing check wins, in the reference's order."""
    L = oracle.lib()
    na, nb, nc = oracle.sz([4, 5, 6]), oracle.sz([3, 5]), oracle.sz([4, 3, 6])
    wa, wc, pia = oracle.sz([1, 4, 20]), oracle.sz([1, 4, 12]), oracle.sz([1, 2, 3])
    x = np.zeros(200)
    a = b = c = x.ctypes.data
    V = L.or_validate
    assert V(2, 3, a, na, wa, pia, b, nb, c, nc, wc) == 0
    assert V(2, 0, a, na, wa, pia, b, nb, c, nc, wc) == 1
    assert V(0, 3, a, na, wa, pia, b, nb, c, nc, wc) == 2
    assert V(4, 3, a, na, wa, pia, b, nb, c, nc, wc) == 2
    assert V(2, 3, None, na, wa, pia, b, nb, c, nc, wc) == 3
    assert V(2, 3, a, na, wa, pia, None, nb, c, nc, wc) == 4
    assert V(2, 3, a, na, wa, pia, b, nb, None, nc, wc) == 5
    assert V(2, 3, a, None, wa, pia, b, nb, c, nc, wc) == 6
    assert V(2, 3, a, na, wa, pia, b, None, c, nc, wc) == 7
    assert V(2, 3, a, na, wa, pia, b, nb, c, None, wc) == 8
    assert V(2, 3, a, na, None, pia, b, nb, c, nc, wc) == 9
    assert V(2, 3, a, na, wa, pia, b, nb, c, nc, None) == 10
    assert V(2, 3, a, na, wa, None, b, nb, c, nc, wc) == 11
    assert V(2, 3, a, na, wa, pia, b, oracle.sz([3, 4]), c, nc, wc) == 12
    assert V(2, 3, a, na, wa, pia, b, oracle.sz([2, 5]), c, nc, wc) == 13
    assert V(2, 3, a, oracle.sz([0, 5, 6]), wa, pia, b, nb, c, nc, wc) == 14
    assert V(2, 3, a, na, wa, pia, b, nb, c, oracle.sz([0, 3, 6]), wc) == 15
    assert V(2, 3, a, na, wa, oracle.sz([1, 1, 3]), b, nb, c, nc, wc) == 16
    assert V(2, 3, a, na, oracle.sz([20, 4, 1]), pia, b, nb, c, nc, wc) == 17
    assert V(2, 3, a, na, wa, pia, b, nb, c, nc, oracle.sz([12, 4, 1])) == 18


def test_planted_transposed_b_is_caught(oracle):
    """A plausible mistake -- reading B with the other orientation -- fails the closed form."""
    na, pia, q, m = [2, 4, 2], [1, 2, 3], 2, 8
    a, wa = rc.ttm_init(q, na, pia)
    b, pib = rc.make_b(m, na[q - 1], "jl", colmajor=False)
    nc = [2, m, 2]
    wc = rc.dense_strides(nc, pia)
    wrong_pib = [1, 2]  # claims column-major storage for a row-major B
    c, _ = oracle.ttm_reference(q, a, na, wa, pia, b, [m, na[q - 1]], wrong_pib, nc, wc)
    assert not np.array_equal(rc.as_nd(c, nc, wc), rc.closed_form(q, na, pia, m, "jl"))
    # and the ones-only B of the reference's own tests would not have caught it
    b1, _ = rc.make_b(m, na[q - 1], "ones", colmajor=False)
    c1, _ = oracle.ttm_reference(q, a, na, wa, pia, b1, [m, na[q - 1]], wrong_pib, nc, wc)
    assert np.array_equal(rc.as_nd(c1, nc, wc), rc.closed_form(q, na, pia, m, "ones"))


def test_generators_are_seeded(oracle):
    x = oracle.fill_uniform(1000, 0x77140001)
    y = oracle.fill_uniform(1000, 0x77140001)
    assert np.array_equal(x, y) and x.min() >= -1 and x.max() <= 1 and abs(x.mean()) < 0.1
    g = np.abs(oracle.fill_loguniform(10000, 5))
    assert g.min() >= 1e-3 * (1 - 1e-6) and g.max() <= 1e3 * (1 + 1e-6)
