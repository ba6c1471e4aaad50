"""Restatement of the reference test-suite's generators and closed-form answers.

* ``generate_shapes`` / ``generate_permutations``: test/include/gtest_aux.h:248-299
* ``ttm_init``: A numbered 1, 2, 3, ... consecutively along mode q, fibres
  enumerated in layout order: test/src/gtest_tlib_ttm.cpp:27-70
* ``closed_form``: with B = ones every C entry is T(a_last) - T(a_first - 1),
  T(n) = n(n+1)/2: test/src/gtest_tlib_ttm.cpp:81-134.  Extended here with a
  B that depends on both indices, B(j, l) = (j + 1) + 100 (l + 1), whose
  answer (j+1) S1 + 100 S2 is still closed form and pins B's orientation
  (the reference tests use B = ones only).
"""
from __future__ import annotations

import itertools

import numpy as np


def generate_shapes(start, dims):
    """All shapes whose r-th extent runs over start[r] * 2**j, j < dims[r]."""
    axes = [[start[r] * 2 ** j for j in range(dims[r])] for r in range(len(start))]
    return [list(s) for s in itertools.product(*axes)]


def generate_permutations(p):
    return [list(x) for x in itertools.permutations(range(1, p + 1))]


def dense_strides(n, pi):
    w = [0] * len(n)
    s = 1
    for mode in pi:
        w[mode - 1] = s
        s *= n[mode - 1]
    return w


def ttm_init(q, na, pia):
    """Flat A (dense in layout pia) numbered along mode q; returns (a, wa)."""
    p = len(na)
    wa = dense_strides(na, pia)
    # fibre number: mixed radix over the modes other than q, fastest in layout
    # order first (the reference's recursion runs r = p .. 1 with q innermost)
    fib = np.zeros(na, dtype=np.float64)
    radix = 1
    for mode in pia:
        if mode == q:
            continue
        shape = [1] * p
        shape[mode - 1] = na[mode - 1]
        fib = fib + radix * np.arange(na[mode - 1], dtype=np.float64).reshape(shape)
        radix *= na[mode - 1]
    shape = [1] * p
    shape[q - 1] = na[q - 1]
    A = fib * na[q - 1] + np.arange(na[q - 1], dtype=np.float64).reshape(shape) + 1
    a = np.zeros(int(np.prod(na)), dtype=np.float64)
    as_nd(a, na, wa)[...] = A
    return a, wa


def as_nd(flat, n, w):
    """Strided view of a flat buffer as an nd-array indexed in natural mode order."""
    item = flat.dtype.itemsize
    return np.lib.stride_tricks.as_strided(flat, shape=tuple(n), strides=tuple(s * item for s in w))


def closed_form(q, na, pia, m, b_kind):
    """Expected C (natural-order nd-array, float64) for ttm_init's A.

    b_kind 'ones': B = 1;  'jl': B(j, l) = (j+1) + 100 (l+1).
    """
    a, wa = ttm_init(q, na, pia)
    A = as_nd(a, na, wa)
    first = np.take(A, [0], axis=q - 1)            # a_first of each fibre
    last = np.take(A, [na[q - 1] - 1], axis=q - 1)  # a_last
    T = lambda n: n * (n + 1) / 2.0
    s1 = T(last) - T(first - 1)                     # sum of the fibre
    if b_kind == "ones":
        return np.repeat(s1, m, axis=q - 1)
    nq = na[q - 1]
    # sum_l (first + l)(l + 1) = first*T(nq) + sum_l l(l+1)
    s2 = first * T(nq) + (nq - 1) * nq * (nq + 1) / 3.0
    shape = [1] * len(na)
    shape[q - 1] = m
    jj = (np.arange(m, dtype=np.float64) + 1).reshape(shape)
    return jj * s1 + 100.0 * s2


def make_b(m, nq, b_kind, colmajor, dtype=np.float64):
    """Flat B in column- (pib={1,2}) or row-major (pib={2,1}) storage."""
    if b_kind == "ones":
        B = np.ones((m, nq))
    else:
        B = (np.arange(m)[:, None] + 1) + 100.0 * (np.arange(nq)[None, :] + 1)
    flat = B.T.reshape(-1) if colmajor else B.reshape(-1)
    return flat.astype(dtype), ([1, 2] if colmajor else [2, 1])


def closed_form_cases(orders=(2, 3, 4), start=2, steps=3, b_kinds=("ones", "jl"), p_filter=None):
    """Yield (na, pia, q, colmajor, b_kind) over the reference's sweep."""
    for p in orders:
        for na in generate_shapes([start] * p, [steps] * p):
            for pia in generate_permutations(p):
                if p_filter and not p_filter(na, pia):
                    continue
                for q in range(1, p + 1):
                    for colmajor in (False, True):
                        for bk in b_kinds:
                            yield na, pia, q, colmajor, bk


def normwise_error(c, c_ref, a, b):
    """||C - C_ref||_F / (||A||_F ||B||_F), all in float64."""
    c = np.asarray(c, dtype=np.float64).reshape(-1)
    c_ref = np.asarray(c_ref, dtype=np.float64).reshape(-1)
    den = np.linalg.norm(np.asarray(a, np.float64)) * np.linalg.norm(np.asarray(b, np.float64))
    return float(np.linalg.norm(c - c_ref) / den) if den else float(np.linalg.norm(c - c_ref))


def eps(dtype):
    return float(np.finfo(dtype).eps)
