"""Validation of batched multi-TTM chain descriptors (no GPU: every call fails before device use)."""
import ctypes

import pytest


def _desc(p=3, n=(4, 5, 6), pi=(1, 2, 3), r=(2, 3, 4), a=0x1000, c=0x2000, factors=True):
    from ttm_amd.chain import ChainDesc
    d = ChainDesc()
    d.p = p
    for k in range(min(p, 8)):
        d.n[k] = n[k] if k < len(n) else 1
        d.pi[k] = pi[k] if k < len(pi) else 0
        d.r[k] = r[k] if k < len(r) else 1
        d.u[k] = 0x3000 + 64 * k if factors else None
    d.a = a
    d.c = c
    return d


@pytest.mark.parametrize("kwargs,status", [
    (dict(p=0), 25), (dict(p=9), 25), (dict(a=None), 25), (dict(c=None), 25),
    (dict(pi=(1, 1, 3)), 16), (dict(pi=(0, 2, 3)), 16), (dict(pi=(1, 2, 4)), 16),
    (dict(n=(4, 0, 6)), 14), (dict(r=(2, 0, 4)), 15),
])
def test_invalid_descriptors(ttmlib, kwargs, status):
    from ttm_amd import chain
    L = chain._bind()
    descs = (chain.ChainDesc * 1)(_desc(**kwargs))
    h = ctypes.c_void_p()
    st = L.ttm_chain_plan_create(1, descs, 0, 0, ctypes.byref(h))
    assert st == status
    assert not h.value
    assert L.ttm_last_error().decode().startswith("Error in tlib::tensor_times_matrix")


def test_precision_rules(ttmlib):
    from ttm_amd import chain
    L = chain._bind()
    descs = (chain.ChainDesc * 1)(_desc())
    h = ctypes.c_void_p()
    assert L.ttm_chain_plan_create(1, descs, 1, 1, ctypes.byref(h)) == 22   # BF16X3 needs float
    assert L.ttm_chain_plan_create(1, descs, 0, 2, ctypes.byref(h)) == 22   # only default / BF16X3
    assert L.ttm_chain_plan_create(1, None, 0, 0, ctypes.byref(h)) == 25
    assert L.ttm_chain_plan_create(1, descs, 0, 0, None) == 25


def test_short_row_schedule_covers_every_unit_once(ttmlib):
    """rows_layout (the short-row kernels' host schedule) on 400 random
    problem lists: contiguous in-order ranges that cover every unit exactly
    once and name the problem of their first unit (tests/cpp/rows_schedule.cpp,
    built against the library; host code only)."""
    import pathlib
    import shutil
    import subprocess
    root = pathlib.Path(__file__).resolve().parents[1]
    gxx = shutil.which("g++")
    if not gxx:
        pytest.skip("no g++")
    build = root / "tests" / "_build"
    build.mkdir(exist_ok=True)
    exe = build / "rows_schedule"
    cmd = [gxx, "-O1", "-std=c++17", "-Wall", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
           f"-I{root / 'include'}", f"-I{root / 'ttm_amd' / 'csrc'}", str(root / "tests" / "cpp" / "rows_schedule.cpp"),
           "-o", str(exe), f"-L{root / 'ttm_amd' / 'lib'}", "-lttm_amd", f"-Wl,-rpath,{root / 'ttm_amd' / 'lib'}"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
    run = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert run.returncode == 0, run.stdout + run.stderr
    assert "0 violations" in run.stdout
