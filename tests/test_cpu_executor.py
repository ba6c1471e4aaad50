"""The host executor (ttm_sttm_cpu / ttm_dttm_cpu, tlib::ttm::cpu) against the
oracle, on the CPU: no GPU is involved.

* the reference's closed-form sweep (test/src/gtest_tlib_ttm.cpp: A numbered
  along q, B of ones, and the extended B(j,l) = (j+1) + 100(l+1)) for every
  one of the 19 tag triples the reference defines, bit-exact;
* random data against the oracle's FP64 brute force, normwise
  ||dC|| / (||A|| ||B||) <= 4 n_q eps;
* the same bits for every triple and thread count (each C element is one
  in-order sum over the contraction);
* an undefined triple is rejected (POLICY) and leaves C untouched, and the
  reference's validation errors come first;
* include/tlib/cpu.h compiled with g++ against the library and run.
"""
import pathlib
import shutil
import subprocess

import numpy as np
import pytest

import refcases as rc

ROOT = pathlib.Path(__file__).resolve().parents[1]

# the 19 triples of include/tlib/detail/ttm.h:405-1474 (ratio variants: the
# parallel_loop_blas ones with a ratio)
TRIPLES = [
    ("sequential", "slice", "none", 0.0), ("parallel_blas", "slice", "none", 0.0),
    ("parallel_taskloop", "slice", "none", 0.0), ("parallel_task", "slice", "none", 0.0),
    ("parallel_blas", "slice", "all", 0.0), ("parallel_loop", "slice", "outer", 0.0),
    ("parallel_loop", "slice", "all", 0.0), ("parallel_loop_blas", "slice", "all", 0.0),
    ("parallel_loop_blas", "slice", "all", 0.25),
    ("sequential", "subtensor", "none", 0.0), ("parallel_blas", "subtensor", "none", 0.0),
    ("parallel_taskloop", "subtensor", "none", 0.0), ("parallel_task", "subtensor", "none", 0.0),
    ("parallel_blas", "subtensor", "all", 0.0), ("parallel_loop", "subtensor", "all", 0.0),
    ("parallel_loop_blas", "subtensor", "all", 0.0), ("parallel_loop_blas", "subtensor", "all", 0.75),
    ("batched_gemm", "subtensor", "all", 0.0), ("combined", "combined", "all", 0.0),
]


def _run(q, a, na, wa, pia, b, nb, pib, nc, wc, triple, threads=4):
    import ttm_amd
    par, sl, fu, ratio = triple
    c = np.full(int(np.prod(nc)), np.nan, dtype=a.dtype)
    ttm_amd.ttm_cpu_raw(q, len(na), a, na, wa, pia, b, nb, pib, c, nc, wc, parallel=par, slicing=sl, fusion=fu,
                        threads=threads, ratio=ratio)
    return c


@pytest.mark.parametrize("triple", TRIPLES, ids=["-".join(map(str, t)) for t in TRIPLES])
def test_closed_form_sweep_every_triple(ttmlib, triple):
    n = 0
    for dtype, kinds in ((np.float64, ("ones", "jl")), (np.float32, ("ones",))):
        for na, pia, q, colmajor, bk in rc.closed_form_cases(orders=(2, 3, 4), b_kinds=kinds):
            if n % 7:  # a seventh of the sweep per triple keeps the suite quick
                n += 1
                continue
            m = 2 * na[q - 1]
            a, wa = rc.ttm_init(q, na, pia)
            b, pib = rc.make_b(m, na[q - 1], bk, colmajor, dtype)
            nc = list(na)
            nc[q - 1] = m
            wc = rc.dense_strides(nc, pia)
            c = _run(q, a.astype(dtype), na, wa, pia, b, [m, na[q - 1]], pib, nc, wc, triple)
            assert np.array_equal(rc.as_nd(c, nc, wc), rc.closed_form(q, na, pia, m, bk)), (triple, na, pia, q, bk)
            n += 1
    assert n > 1000


def test_random_data_vs_oracle_and_identical_bits(ttmlib, oracle):
    rng = np.random.default_rng(61)
    cases = [([37, 41, 29], [1, 2, 3], 2, 19), ([37, 41, 29], [1, 2, 3], 1, 70), ([37, 41, 29], [3, 2, 1], 3, 5),
             ([3, 700, 9], [1, 2, 3], 2, 33), ([2, 5, 300], [1, 2, 3], 3, 130), ([64, 48, 32], [1, 2, 3], 2, 40),
             ([1500, 3], [1, 2], 1, 16), ([5000], [1], 1, 7), ([12, 7, 6, 9], [3, 1, 4, 2], 4, 11)]
    for dtype in (np.float32, np.float64):
        for na, pia, q, m in cases:
            for colmajor in (False, True):
                wa = rc.dense_strides(na, pia)
                a = rng.uniform(-1, 1, int(np.prod(na))).astype(dtype)
                Bm = rng.uniform(-1, 1, (m, na[q - 1])).astype(dtype)
                b = (Bm.T if colmajor else Bm).reshape(-1).copy()
                nc = list(na)
                nc[q - 1] = m
                wc = rc.dense_strides(nc, pia)
                nb, pib = [m, na[q - 1]], ([1, 2] if colmajor else [2, 1])
                c64 = oracle.ttm_brute_f64(q, a, na, wa, b, m, colmajor, nc, wc)
                first = None
                for triple in TRIPLES:
                    for threads in ((1, 3, 8) if triple[0] in ("parallel_loop", "combined") else (5,)):
                        c = _run(q, a, na, wa, pia, b, nb, pib, nc, wc, triple, threads)
                        if first is None:
                            first = c
                            err = rc.normwise_error(c, c64, a, b)
                            assert err <= 4 * na[q - 1] * rc.eps(dtype), (na, pia, q, m, err)
                        assert np.array_equal(c, first), (triple, threads, na, pia, q, m)


def test_undefined_triple_and_validation(ttmlib):
    import ttm_amd
    na, pia, q, m = [4, 5, 6], [1, 2, 3], 2, 3
    wa = rc.dense_strides(na, pia)
    a = np.ones(120)
    b = np.ones(15)
    nc = [4, 3, 6]
    wc = rc.dense_strides(nc, pia)
    c = np.full(72, np.nan)
    for par, sl, fu in (("batched_gemm", "slice", "all"), ("sequential", "subtensor", "all"),
                        ("combined", "subtensor", "all"), ("parallel_loop", "slice", "none")):
        with pytest.raises(ttm_amd.TTMError) as e:
            ttm_amd.ttm_cpu_raw(q, 3, a, na, wa, pia, b, [m, 5], [1, 2], c, nc, wc, parallel=par, slicing=sl,
                                fusion=fu)
        assert e.value.status == ttm_amd.Status.POLICY and "not yet defined" in str(e.value)
    assert np.isnan(c).all()
    # the reference's checks come first (include/tlib/ttm.h:65-90)
    with pytest.raises(ttm_amd.TTMError) as e:
        ttm_amd.ttm_cpu_raw(q, 3, a, na, wa, pia, b, [m, 4], [1, 2], c, nc, wc, parallel="batched_gemm",
                            slicing="slice", fusion="all")
    assert e.value.status == ttm_amd.Status.CONTRACTION_DIM
    assert np.isnan(c).all()


CPP = r"""
#include <tlib/cpu.h>
#include <cmath>
#include <cstdio>
#include <vector>
int main()
{
    namespace tt = tlib::ttm;
    // A 4x3x2 numbered along mode 2, B of ones: the reference's closed form
    auto A = tt::tensor<double>({4, 3, 2}, {1, 2, 3});
    for (std::size_t i = 0; i < A.data().size(); ++i) A[i] = double(i % 12 / 4 + 1);
    auto B = tt::tensor<double>({5, 3}, {1, 2});
    for (auto &x : B.data()) x = 1.0;
    auto C1 = tt::cpu::ttm(2u, A, B, tt::parallel_policy::parallel_taskloop, tt::slicing_policy::subtensor,
                           tt::fusion_policy::none);
    auto C2 = tt::cpu::ttm(2u, A, B, tt::parallel_policy::combined, tt::slicing_policy::combined,
                           tt::fusion_policy::all);
    std::vector<double> C3(C1.data().size(), -1.0);
    tt::cpu::ttm(tt::parallel_policy::parallel_loop_blas, tt::slicing_policy::slice, tt::fusion_policy::all, 2u, 3u,
                 0.5, A.data().data(), A.shape().data(), A.strides().data(), A.layout().data(), B.data().data(),
                 B.shape().data(), B.layout().data(), C3.data(), C1.shape().data(), C1.strides().data());
    // an undefined triple prints the reference's message and leaves C alone
    std::vector<double> C4(C1.data().size(), -1.0);
    tt::cpu::ttm(tt::parallel_policy::batched_gemm, tt::slicing_policy::slice, tt::fusion_policy::all, 2u, 3u,
                 A.data().data(), A.shape().data(), A.strides().data(), A.layout().data(), B.data().data(),
                 B.shape().data(), B.layout().data(), C4.data(), C1.shape().data(), C1.strides().data());
    int bad = 0;
    for (std::size_t i = 0; i < C1.data().size(); ++i) {
        if (C1[i] != 6.0 || C2[i] != 6.0 || C3[i] != 6.0 || C4[i] != -1.0) ++bad;
    }
    try {
        auto Bw = tt::tensor<double>({5, 4}, {1, 2});
        (void)tt::cpu::ttm(2u, A, Bw, tt::parallel_policy::sequential, tt::slicing_policy::slice,
                           tt::fusion_policy::none);
        ++bad;
    } catch (std::runtime_error const &e) {
        std::printf("%s\n", e.what());
    }
    std::printf("bad %d\n", bad);
    return bad;
}
"""


def test_cpp_header_runs_on_the_host(ttmlib, tmp_path):
    gxx = shutil.which("g++")
    if not gxx:
        pytest.skip("no g++")
    src = tmp_path / "cpu.cpp"
    src.write_text(CPP)
    exe = tmp_path / "cpu"
    libdir = ttmlib.LIB_PATH.parent
    build = subprocess.run([gxx, "-std=c++17", "-O1", f"-I{ROOT / 'include'}", str(src), f"-L{libdir}",
                            "-lttm_amd", f"-Wl,-rpath,{libdir}", "-o", str(exe)], capture_output=True, text=True)
    assert build.returncode == 0, build.stderr
    run = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert run.returncode == 0, run.stdout + run.stderr
    assert "This type of ttm is not yet defined" in run.stdout
    assert "contraction dimension of A and B are not equal" in run.stdout
    assert "bad 0" in run.stdout
