"""ctypes binding of the CPU oracle (oracle/libttm_oracle.so) -- TEST INFRASTRUCTURE ONLY.

Imported only by tests/, ``__graft_entry__.smoke()`` and the cpu_baseline
leg of bench.py, as the reference result / CPU baseline; the product never uses it.
Every function states the reference lines its C counterpart follows (see
ttm_oracle.c).
"""
from __future__ import annotations

import ctypes
import glob
import os
import pathlib
import subprocess

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
LIB = HERE / "libttm_oracle.so"
REF_LIB = HERE / "_ref" / "libttm_ref_meta.so"
REF_INC = pathlib.Path("/root/reference/include")

_sp = ctypes.POINTER(ctypes.c_size_t)
_vp = ctypes.c_void_p
_lib = None
_ref = None


def build(with_ref: bool = True) -> None:
    """Compile the oracle (and, where the reference sources exist, oracle/_ref)."""
    subprocess.run(["make", "-s", "-C", str(HERE), "all"], check=True)
    if with_ref and REF_INC.is_dir():
        subprocess.run(["make", "-s", "-C", str(HERE), "ref", f"REF_INC={REF_INC}"], check=True)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build(with_ref=False)
        L = ctypes.CDLL(str(LIB))
        ttm_args = [ctypes.c_uint, ctypes.c_uint, _vp, _sp, _sp, _sp, _vp, _sp, _sp, _vp, _sp, _sp,
                    ctypes.c_int, ctypes.c_uint, ctypes.POINTER(ctypes.c_int)]
        sig = {
            "or_is_valid_shape": (ctypes.c_int, [_sp, ctypes.c_size_t]),
            "or_is_scalar": (ctypes.c_int, [_sp, ctypes.c_size_t]),
            "or_is_vector": (ctypes.c_int, [_sp, ctypes.c_size_t]),
            "or_is_matrix": (ctypes.c_int, [_sp, ctypes.c_size_t]),
            "or_is_tensor": (ctypes.c_int, [_sp, ctypes.c_size_t]),
            "or_is_valid_layout": (ctypes.c_int, [_sp, ctypes.c_size_t]),
            "or_compute_k_order_layout": (ctypes.c_int, [_sp, ctypes.c_size_t, ctypes.c_size_t]),
            "or_is_first_order": (ctypes.c_int, [_sp, ctypes.c_size_t]),
            "or_is_last_order": (ctypes.c_int, [_sp, ctypes.c_size_t]),
            "or_inverse_mode": (ctypes.c_size_t, [_sp, ctypes.c_size_t, ctypes.c_size_t]),
            "or_compute_strides": (ctypes.c_int, [_sp, _sp, ctypes.c_size_t, _sp]),
            "or_is_valid_strides": (ctypes.c_int, [_sp, ctypes.c_size_t, _sp]),
            "or_is_case": (ctypes.c_int, [ctypes.c_uint, ctypes.c_uint, ctypes.c_size_t, _sp]),
            "or_product": (ctypes.c_size_t, [_sp, _sp, ctypes.c_uint, ctypes.c_uint]),
            "or_cores": (ctypes.c_uint, []),
            "or_validate": (ctypes.c_int, [ctypes.c_uint, ctypes.c_uint, _vp, _sp, _sp, _sp, _vp, _sp,
                                           _vp, _sp, _sp]),
            "or_ttm_s": (ctypes.c_int, ttm_args),
            "or_ttm_d": (ctypes.c_int, ttm_args),
            "or_ttm_brute_f64": (None, [ctypes.c_uint, ctypes.c_uint, _vp, ctypes.c_int, _sp, _sp, _vp,
                                        ctypes.c_int, ctypes.c_size_t, ctypes.c_int, _vp, _sp]),
            "or_blas_load": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int]),
            "or_blas_name": (ctypes.c_char_p, []),
            "or_fill_uniform_s": (None, [_vp, ctypes.c_size_t, ctypes.c_uint64]),
            "or_fill_uniform_d": (None, [_vp, ctypes.c_size_t, ctypes.c_uint64]),
            "or_fill_loguniform_s": (None, [_vp, ctypes.c_size_t, ctypes.c_uint64]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def ref_lib():
    """The reference's metadata headers compiled under oracle/_ref, or None."""
    global _ref
    if _ref is None and REF_LIB.exists():
        R = ctypes.CDLL(str(REF_LIB))
        for name in ("ref_is_valid_shape", "ref_is_scalar", "ref_is_vector", "ref_is_matrix",
                     "ref_is_tensor", "ref_is_valid_layout", "ref_is_first_order", "ref_is_last_order"):
            getattr(R, name).restype = ctypes.c_int
            getattr(R, name).argtypes = [_sp, ctypes.c_size_t]
        R.ref_compute_k_order_layout.restype = ctypes.c_int
        R.ref_compute_k_order_layout.argtypes = [_sp, ctypes.c_size_t, ctypes.c_size_t]
        R.ref_inverse_mode.restype = ctypes.c_size_t
        R.ref_inverse_mode.argtypes = [_sp, ctypes.c_size_t, ctypes.c_size_t]
        R.ref_compute_strides.restype = ctypes.c_int
        R.ref_compute_strides.argtypes = [_sp, _sp, ctypes.c_size_t, _sp]
        R.ref_is_valid_strides.restype = ctypes.c_int
        R.ref_is_valid_strides.argtypes = [_sp, ctypes.c_size_t, _sp]
        R.ref_is_case.restype = ctypes.c_int
        R.ref_is_case.argtypes = [ctypes.c_uint, ctypes.c_uint, ctypes.c_size_t, _sp]
        R.ref_tensor_strides.restype = ctypes.c_int
        R.ref_tensor_strides.argtypes = [_sp, _sp, ctypes.c_size_t, _sp]
        _ref = R
    return _ref


def sz(values):
    vals = [int(v) for v in values]
    return (ctypes.c_size_t * max(1, len(vals)))(*vals)


def _ptr(x: np.ndarray):
    return ctypes.c_void_p(x.ctypes.data)


# --- metadata ---------------------------------------------------------------

def compute_strides(n, pi):
    """include/tlib/detail/strides.h:31-57 (None where the reference throws)."""
    w = sz([0] * len(n))
    if lib().or_compute_strides(sz(n), sz(pi), len(n), w) != 0:
        return None
    return list(w)[: len(n)]


def k_order_layout(p, k):
    """include/tlib/detail/layout.h:57-84."""
    pi = sz([0] * p)
    lib().or_compute_k_order_layout(pi, p, k)
    return list(pi)[:p]


def inverse_mode(pi, mode):
    """include/tlib/detail/layout.h:120-144 (0 where the reference throws)."""
    return lib().or_inverse_mode(sz(pi), len(pi), mode)


def is_case(case_nr, p, q, pi):
    return bool(lib().or_is_case(case_nr, p, q, sz(pi)))


def cores():
    """include/tlib/detail/ttm.h:60-92."""
    return lib().or_cores()


# --- contraction --------------------------------------------------------------

def ttm_reference(q, a, na, wa, pia, b, nb, pib, nc, wc, *, use_blas=False, cores=0):
    """The reference's `combined` algorithm (include/tlib/detail/ttm.h:1408-1474).

    a, b are flat float32/float64 arrays; returns (c_flat, branch) with c in
    C's dense layout.  Raises ValueError(check_number) when validation fails.
    """
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    dt = a.dtype
    c = np.zeros(int(np.prod(nc)), dtype=dt)
    br = ctypes.c_int(-1)
    fn = lib().or_ttm_s if dt == np.float32 else lib().or_ttm_d
    st = fn(q, len(na), _ptr(a), sz(na), sz(wa), sz(pia), _ptr(b), sz(nb), sz(pib), _ptr(c), sz(nc), sz(wc),
            int(use_blas), int(cores), ctypes.byref(br))
    if st:
        raise ValueError(st)
    return c, br.value


def ttm_brute_f64(q, a, na, wa, b, m, b_colmajor, nc, wc):
    """FP64 brute force of the definition (long-double accumulation) over any strides.

    a, b flat float32/float64 arrays; returns a flat float64 C of size prod(nc)
    laid out by the element strides wc.
    """
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    size_c = 1 + sum((n - 1) * w for n, w in zip(nc, wc))
    c = np.zeros(size_c, dtype=np.float64)
    lib().or_ttm_brute_f64(q, len(na), _ptr(a), int(a.dtype == np.float64), sz(na), sz(wa), _ptr(b),
                           int(b.dtype == np.float64), int(m), int(bool(b_colmajor)), _ptr(c), sz(wc))
    return c


def fill_uniform(n, seed, dtype=np.float32):
    """SplitMix64 uniform[-1,1] (shared with the bench)."""
    x = np.empty(int(n), dtype=dtype)
    (lib().or_fill_uniform_s if dtype == np.float32 else lib().or_fill_uniform_d)(_ptr(x), x.size, seed)
    return x


def fill_loguniform(n, seed):
    """SplitMix64 sign-random log-uniform magnitudes in [1e-3, 1e3] (float32)."""
    x = np.empty(int(n), dtype=np.float32)
    lib().or_fill_loguniform_s(_ptr(x), x.size, seed)
    return x


def bind_blas() -> str:
    """Bind a CBLAS for the timed CPU baseline: OpenBLAS, NumPy's bundled
    scipy-openblas (ILP64), or MKL; returns its path or 'naive'."""
    L = lib()
    cands = []
    for path in ("libopenblas.so.0", "libopenblas.so"):
        cands.append((path, 0))
    try:
        import numpy
        libs = pathlib.Path(numpy.__file__).resolve().parent.parent / "numpy.libs"
        for path in sorted(glob.glob(str(libs / "libscipy_openblas64_*.so"))):
            cands.append((path, 1))
    except Exception:  # pragma: no cover
        pass
    cands.append(("/opt/conda/lib/libmkl_rt.so", 0))
    for path, kind in cands:
        if kind == 0 and "/" in path and not os.path.exists(path):
            continue
        if L.or_blas_load(path.encode(), kind) == 0:
            return L.or_blas_name().decode()
    return "naive"
