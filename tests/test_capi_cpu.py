"""CPU-side checks of the product: the C ABI library, its validation and the drop-in headers.

No kernel runs here: every call below fails validation before any HIP call,
or only plans.
"""
import ctypes
import pathlib
import re
import shutil
import subprocess

import numpy as np
import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "ttm_amd.h"
BUILD = ROOT / "tests" / "_build"
REF_INCLUDE = pathlib.Path("/root/reference/include")
REF_EXAMPLES = pathlib.Path("/root/reference/example")

REFERENCE_MESSAGES = {
    1: "input tensor order should be greater zero.",
    2: "contraction mode should be greater zero or less than or equal to p.",
    3: "pointer to input tensor A should not be zero.",
    4: "pointer to input vector B should not be zero.",
    5: "pointer to output tensor C should not be zero.",
    6: "pointer to input tensor shape vector na should not be zero.",
    7: "pointer to input vector shape vector nb should not be zero.",
    8: "pointer to output tensor shape vector nc should not be zero.",
    9: "pointer to input tensor stride vector wa should not be zero.",
    10: "pointer to output tensor stride vector wc should not be zero.",
    11: "pointer to input tensor permutation vector pia should not be zero.",
    12: "contraction dimension of A and B are not equal.",
    13: "free dimension of C and B are not equal.",
    14: "shape vector of A is not valid.",
    15: "shape vector of C is not valid.",
    16: "layout vector of A is not valid.",
    17: "stride vector of A is not valid.",
    18: "stride vector of C is not valid.",
}


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ttm_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(ttmlib):
    L = ttmlib.lib()
    names = declared_functions()
    assert len(names) >= 12
    for name in names:
        assert hasattr(L, name), name
    nm = shutil.which("nm")
    if nm:
        out = subprocess.run([nm, "-D", "--defined-only", str(ttmlib.LIB_PATH)], capture_output=True, text=True).stdout
        exported = set(re.findall(r" T (\w+)", out))
        assert set(names) <= exported


def test_library_targets_gfx950(ttmlib):
    bundler = pathlib.Path("/opt/rocm/llvm/bin/clang-offload-bundler")
    if not bundler.exists():
        pytest.skip("no offload bundler")
    out = subprocess.run([str(bundler), "--list", "--type=o", f"--input={ttmlib.LIB_PATH}"],
                         capture_output=True, text=True)
    # the fat binary may be embedded in .hip_fatbin; fall back to searching the bytes
    blob = ttmlib.LIB_PATH.read_bytes()
    assert b"gfx950" in blob or "gfx950" in out.stdout


def _args(q=2, p=3, a=True, b=True, c=True, na=(4, 5, 6), wa=(1, 4, 20), pia=(1, 2, 3), nb=(3, 5),
          pib=(1, 2), nc=(4, 3, 6), wc=(1, 4, 12), dtype=np.float32):
    from ttm_amd._lib import sizes
    buf = np.zeros(1024, dtype=dtype)
    ptr = buf.ctypes.data
    keep = [buf]

    def arr(v):
        if v is None:
            return None
        x = sizes(v)
        keep.append(x)
        return x

    return keep, [q, p, ptr if a else None, arr(na), arr(wa), arr(pia), ptr if b else None, arr(nb),
                  arr(pib), ptr if c else None, arr(nc), arr(wc)]


@pytest.mark.parametrize("entry", ["ttm_sttm_host", "ttm_dttm_host", "ttm_sttm", "ttm_dttm", "ttm_sttm_ex"])
def test_validation_statuses_and_messages(ttmlib, entry):
    L = ttmlib.lib()
    fn = getattr(L, entry)
    dtype = np.float64 if entry.startswith("ttm_d") else np.float32
    tail = {"ttm_sttm": [None], "ttm_dttm": [None], "ttm_sttm_ex": [0, None]}.get(entry, [])
    cases = {
        1: dict(p=0), 2: dict(q=0), 3: dict(a=False), 4: dict(b=False), 5: dict(c=False),
        6: dict(na=None), 7: dict(nb=None), 8: dict(nc=None), 9: dict(wa=None), 10: dict(wc=None),
        11: dict(pia=None), 12: dict(nb=(3, 4)), 13: dict(nb=(2, 5)), 14: dict(na=(0, 5, 6)),
        15: dict(nc=(4, 3, 0)), 16: dict(pia=(1, 1, 3)), 17: dict(wa=(20, 4, 1)), 18: dict(wc=(12, 4, 1)),
        19: dict(pib=None), 20: dict(wa=(1, 4, 21)), 21: dict(nc=(4, 3, 7), wc=(1, 4, 12)),
    }
    for status, kw in cases.items():
        keep, args = _args(dtype=dtype, **kw)
        st = fn(*args, *tail)
        assert st == status, (entry, status, st)
        msg = L.ttm_last_error().decode()
        assert msg.startswith("Error in tlib::tensor_times_matrix: ")
        if status in REFERENCE_MESSAGES:
            assert msg == "Error in tlib::tensor_times_matrix: " + REFERENCE_MESSAGES[status]
        assert L.ttm_status_message(status).decode() == msg.split(" (")[0]


def test_q_beyond_p_and_order_checks(ttmlib):
    L = ttmlib.lib()
    keep, args = _args(q=4)
    assert L.ttm_sttm_host(*args) == 2
    keep, args = _args(p=0, q=0)
    assert L.ttm_sttm_host(*args) == 1  # p is checked first


def test_double_precision_variants_rejected(ttmlib):
    import ttm_amd
    x = np.zeros((4, 5, 6))
    with pytest.raises(ttm_amd.TTMError) as e:
        ttm_amd.ttm(2, x, np.zeros((3, 5)), precision="bf16x3")
    assert e.value.status == ttm_amd.Status.PRECISION


def test_device_memory_entry_points_without_device_work(ttmlib):
    """The paths of ttm_device_* that touch no device: empty sizes, null pointers."""
    import ctypes
    import ttm_amd
    L = ttmlib.lib()
    p = ctypes.c_void_p(12345)
    assert L.ttm_device_alloc(0, ctypes.byref(p)) == 0 and not p.value  # nothing allocated, NULL returned
    assert L.ttm_device_alloc(16, None) == ttm_amd.Status.ARGUMENT
    assert b"null" in L.ttm_last_error()
    assert L.ttm_device_free(None) == 0
    assert L.ttm_device_copy(None, None, 0) == 0
    assert L.ttm_device_copy(None, ctypes.c_void_p(8), 4) == ttm_amd.Status.ARGUMENT
    assert L.ttm_device_zero(None, 0) == 0
    assert L.ttm_device_zero(None, 4) == ttm_amd.Status.ARGUMENT
    assert L.ttm_last_error() != b""


def test_tucker_rank_checks_before_any_device_work():
    import torch
    from ttm_amd.tucker import factors_of
    A = torch.zeros(4, 5, 6, dtype=torch.float64)
    for bad in ([0, 2, 2], [2, 6, 2], [2, 2, 7]):
        with pytest.raises(ValueError):
            factors_of(A, bad)


def test_plan_reference_configs(ttmlib):
    """Kernel families the planner picks for the A 1024x1024x512 sweep (no GPU needed)."""
    import ttm_amd
    na = [1024, 1024, 512]
    for pia in ([1, 2, 3], [3, 2, 1]):
        for q in (1, 2, 3):
            for m in (16, 1024):
                nb = [m, na[q - 1]]
                qh = pia.index(q) + 1
                for dbl in (False, True):
                    k = ttm_amd.plan_kernel(q, na, pia, nb, [1, 2], is_double=dbl)
                    assert k, (pia, q, m)
                    sym = ttm_amd.plan_kernel_symbol(q, na, pia, nb, [1, 2], is_double=dbl)
                    if m == 16 and qh > 1:
                        # NN view (w % vec == 0), one row block: the streaming kernel
                        assert k == "nn_stream", (pia, q, m, k)
                        assert sym.startswith("k_nn_stream<" + ("double" if dbl else "float") + ", 1, 6, ")
                        continue
                    if m == 16:
                        # TN view (q fastest) with all of B in LDS: 8 waves, or 3 for 513-piece B rows
                        nq = na[q - 1]
                        spb = {(False, 1024): "8, 2, 257", (False, 512): "8, 2, 129",
                               (True, 1024): "3, 2, 513", (True, 512): "8, 2, 257"}[(dbl, nq)]
                        assert k == "tn_stream", (pia, q, m, k)
                        assert sym == "k_tn_lds<" + ("double" if dbl else "float") + ", " + spb + ">", sym
                        continue
                    assert k == "gemm", (pia, q, m, k)
                    assert sym.startswith("k_gemm<" + ("double" if dbl else "float"))
                    assert sym.endswith(("true, true, false>" if qh == 1 else "false, true, false>")), sym


def test_plan_odd_extents_and_underfilled_grids(ttmlib):
    """Odd extents take the tile kernels (padded NN slabs, element-aligned
    rows), GEMV-like shapes the skinny kernel, and long contractions with few
    blocks a contraction split (no GPU needed)."""
    import ttm_amd
    S, K, N = ttm_amd.plan_kernel_symbol, ttm_amd.plan_kernel, ttm_amd.plan_splits
    for dbl in (False, True):
        t = "double" if dbl else "float"
        for m in (16, 1024):
            # w = 999: NN tile with padded slabs
            assert S(2, [999, 1000, 500], [1, 2, 3], [m, 1000], [1, 2], dbl).endswith("false, true, true>")
            # n_q = 1023 / 513 with q fastest: TN tile on element-aligned rows
            assert S(1, [1023, 1024, 512], [1, 2, 3], [m, 1023], [1, 2], dbl).startswith(f"k_gemm<{t}")
        assert S(3, [1024, 1024, 513], [3, 2, 1], [16, 513], [1, 2], dbl).startswith(f"k_gemm<{t}")
        assert K(2, [3, 5, 7], [1, 2, 3], [4, 5], [1, 2], dbl) == "gemm"
        # p = 1 (the reference's GEMV case) and [2^20, 4] with q = 1: skinny, split
        for m in (1, 16):
            assert K(1, [1 << 24], [1], [m, 1 << 24], [2, 1], dbl) == "skinny"
            assert N(1, [1 << 24], [1], [m, 1 << 24], [2, 1], dbl) > 64
        assert K(1, [1 << 20, 4], [1, 2], [1024, 1 << 20], [1, 2], dbl) == "skinny"
        assert N(1, [1 << 20, 4], [1, 2], [1024, 1 << 20], [1, 2], dbl) > 8
        # [16, 16, 2^20] with q = 3: 16 (or 2) blocks -> split along the contraction
        for m in (16, 1024):
            assert K(3, [16, 16, 1 << 20], [1, 2, 3], [m, 1 << 20], [1, 2], dbl) == "gemm"
            assert N(3, [16, 16, 1 << 20], [1, 2, 3], [m, 1 << 20], [1, 2], dbl) >= 16
        # the bench shapes fill the chip: no split
        assert N(2, [1024, 1024, 512], [1, 2, 3], [1024, 1024], [2, 1], dbl) == 1


def _gxx():
    gxx = shutil.which("g++")
    if not gxx:
        pytest.skip("no g++")
    return gxx


def test_product_metadata_headers_match_reference(oracle):
    """include/tlib/detail/* compiled into the same shim as oracle/_ref, compared exhaustively."""
    import itertools
    gxx = _gxx()
    BUILD.mkdir(exist_ok=True)
    so = BUILD / "libttm_product_meta.so"
    subprocess.run([gxx, "-O1", "-std=c++17", "-fPIC", "-shared", "-w", f"-I{ROOT / 'include'}",
                    str(ROOT / "oracle" / "ref_meta.cpp"), "-o", str(so)], check=True)
    P = ctypes.CDLL(str(so))
    R = oracle.ref_lib()
    L = oracle.lib()
    sz = oracle.sz
    for name in ("ref_is_valid_shape", "ref_is_scalar", "ref_is_vector", "ref_is_matrix", "ref_is_tensor",
                 "ref_is_valid_layout", "ref_is_first_order", "ref_is_last_order"):
        getattr(P, name).argtypes = [ctypes.POINTER(ctypes.c_size_t), ctypes.c_size_t]
    P.ref_inverse_mode.restype = ctypes.c_size_t
    # against the compiled reference when it exists, else against the (pinned) oracle
    def other(name, *args):
        if R is not None:
            return getattr(R, name)(*args)
        return getattr(L, "or_" + name[4:])(*args)
    for p in range(0, 5):
        for s in itertools.product(range(4), repeat=p):
            for f in ("is_valid_shape", "is_scalar", "is_vector", "is_matrix", "is_tensor"):
                assert getattr(P, "ref_" + f)(sz(s), p) == other("ref_" + f, sz(s), p), (f, s)
    for p in range(1, 5):
        for lay in itertools.product(range(p + 2), repeat=p):
            pl = sz(lay)
            assert P.ref_is_valid_layout(pl, p) == other("ref_is_valid_layout", pl, p)
            for mode in range(0, p + 2):
                assert P.ref_inverse_mode(pl, p, mode) == other("ref_inverse_mode", pl, p, mode)
            for case in range(1, 9):
                for q in range(1, p + 1):
                    assert P.ref_is_case(case, p, q, pl) == other("ref_is_case", case, p, q, pl)
    for p in range(1, 8):
        for k in range(0, p + 3):
            a, b = sz([0] * p), sz([0] * p)
            P.ref_compute_k_order_layout(a, p, k)
            other("ref_compute_k_order_layout", b, p, k)
            assert list(a) == list(b)
    rng = np.random.default_rng(11)
    for p in range(1, 5):
        for s in itertools.product((1, 2, 3), repeat=p):
            for lay in itertools.permutations(range(1, p + 1)):
                w1, w2 = sz([0] * p), sz([0] * p)
                assert P.ref_compute_strides(sz(s), sz(lay), p, w1) == other("ref_compute_strides", sz(s), sz(lay), p, w2)
                assert list(w1) == list(w2)
                w3 = sz([0] * p)
                assert P.ref_tensor_strides(sz(s), sz(lay), p, w3) == 0 and list(w3) == list(w1)
                for _ in range(3):
                    rw = sz(rng.integers(1, 6, size=p).tolist())
                    assert P.ref_is_valid_strides(sz(lay), p, rw) == other("ref_is_valid_strides", sz(lay), p, rw)


def _compile_cpp(src, out, extra=()):
    gxx = _gxx()
    BUILD.mkdir(exist_ok=True)
    cmd = [gxx, "-O1", "-std=c++17", "-Wall", f"-I{ROOT / 'include'}", str(src), "-o", str(out),
           f"-L{ROOT / 'ttm_amd' / 'lib'}", "-lttm_amd", f"-Wl,-rpath,{ROOT / 'ttm_amd' / 'lib'}", *extra]
    res = subprocess.run(cmd, capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
    return out


def test_dropin_header_program_compiles(ttmlib):
    _compile_cpp(ROOT / "tests" / "cpp" / "interfaces.cpp", BUILD / "interfaces")


def test_device_tensor_program_compiles(ttmlib):
    _compile_cpp(ROOT / "tests" / "cpp" / "device_tensor.cpp", BUILD / "device_tensor")


@pytest.mark.skipif(not REF_EXAMPLES.is_dir(), reason="reference examples not available")
@pytest.mark.parametrize("example", ["interface1.cpp", "interface2.cpp", "interface3.cpp"])
def test_reference_examples_compile_unchanged(ttmlib, example):
    """The reference's own example programs build against the drop-in header."""
    _compile_cpp(REF_EXAMPLES / example, BUILD / example.replace(".cpp", ""), extra=["-Wno-unused-variable"])
