"""The Python front end rejects mismatched buffers before the library sees a
pointer: mixed residencies, element types other than A's, buffers smaller
than their extents or not one dense block, and chain descriptors over host
tensors or mixed element types.  No GPU is needed (every case raises before
any device work)."""
import numpy as np
import pytest


def _case(dtype_a=np.float32, dtype_b=None, dtype_c=None):
    na, pia, q, m = [4, 5, 6], [1, 2, 3], 2, 3
    nc = [4, m, 6]
    wa = [1, 4, 20]
    wc = [1, 4, 12]
    a = np.zeros(int(np.prod(na)), dtype=dtype_a)
    b = np.zeros(m * 5, dtype=dtype_b or dtype_a)
    c = np.zeros(int(np.prod(nc)), dtype=dtype_c or dtype_a)
    return q, 3, a, na, wa, pia, b, [m, 5], [2, 1], c, nc, wc


def test_b_and_c_must_have_a_dtype(ttmlib):
    import ttm_amd
    for kw in (dict(dtype_a=np.float64, dtype_b=np.float32), dict(dtype_a=np.float32, dtype_b=np.float64),
               dict(dtype_a=np.float32, dtype_c=np.float64)):
        with pytest.raises(TypeError, match="element type"):
            ttm_amd.ttm_raw(*_case(**kw))


def test_mixed_residency_rejected(ttmlib):
    import ttm_amd
    q, p, a, na, wa, pia, b, nb, pib, c, nc, wc = _case()
    with pytest.raises(TypeError, match="all be"):
        ttm_amd.ttm_raw(q, p, 1 << 40, na, wa, pia, b, nb, pib, c, nc, wc)
    with pytest.raises(TypeError, match="all be"):
        ttm_amd.ttm_raw(q, p, a, na, wa, pia, b, nb, pib, 1 << 40, nc, wc)
    with pytest.raises(TypeError, match="dtype="):
        ttm_amd.ttm_raw(q, p, 1 << 40, na, wa, pia, 1 << 41, nb, pib, 1 << 42, nc, wc)
    with pytest.raises(TypeError, match="unsupported buffer"):
        ttm_amd.ttm_raw(q, p, [0.0] * 120, na, wa, pia, b, nb, pib, c, nc, wc)


def test_short_or_gapped_buffers_rejected(ttmlib):
    import ttm_amd
    q, p, a, na, wa, pia, b, nb, pib, c, nc, wc = _case()
    with pytest.raises(ValueError, match="B holds 14"):
        ttm_amd.ttm_raw(q, p, a, na, wa, pia, b[:14], nb, pib, c, nc, wc)
    with pytest.raises(ValueError, match="C holds"):
        ttm_amd.ttm_raw(q, p, a, na, wa, pia, b, nb, pib, c[:-1], nc, wc)
    big = np.zeros(2 * a.size, dtype=a.dtype)
    with pytest.raises(ValueError, match="dense block"):
        ttm_amd.ttm_raw(q, p, big[::2], na, wa, pia, b, nb, pib, c, nc, wc)


def test_dense_numel_of_permuted_views():
    import torch
    from ttm_amd.api import _dense_numel
    x = np.zeros((3, 4, 5), dtype=np.float32)
    assert _dense_numel(x) == 60
    assert _dense_numel(np.asfortranarray(x)) == 60
    assert _dense_numel(x.transpose(2, 0, 1)) == 60
    assert _dense_numel(x[:, ::2]) is None
    t = torch.zeros(3, 4, 5)
    assert _dense_numel(t.permute(1, 2, 0)) == 60
    assert _dense_numel(t[:, 1:]) is None
    assert _dense_numel(torch.zeros(3, 1, 5)[:, :, :]) == 15


def test_tensor_front_end_out_must_match(ttmlib):
    import ttm_amd
    A = np.zeros((4, 5, 6), dtype=np.float32)
    B = np.zeros((3, 5), dtype=np.float32)
    out = np.zeros((4, 3, 6), dtype=np.float64)
    with pytest.raises(TypeError, match="out must have"):
        ttm_amd.ttm(2, A, B, out=out)


def test_chain_descriptors_need_device_tensors_and_one_dtype(ttmlib):
    import torch
    from ttm_amd.chain import ChainPlan, make_desc
    A = torch.zeros(4, 5, 6)
    C = torch.zeros(2, 5, 6)
    U = torch.zeros(2, 4)
    with pytest.raises(TypeError, match="ROCm device"):
        make_desc(A, [U, None, None], C)
    A64 = torch.zeros(4, 5, 6, dtype=torch.float64)
    C64 = torch.zeros(2, 5, 6, dtype=torch.float64)
    with pytest.raises(TypeError, match="same element type"):
        ChainPlan([(A, [U, None, None], C), (A64, [U.double(), None, None], C64)])
