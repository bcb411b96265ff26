"""rn_tv_sqs_noalloc (allocation-free variant of rn_tv_sqs) against rn_tv_sqs.

Both run the same kernels, so gradient, curvature and the FP64 per-slice costs
must agree bit for bit; two calls in a row, the second on a smaller input, so
that stale partials of the first call in the shared device array would show.
A size beyond the array is refused on the host before any launch.
"""

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

P, ci, cf = ctypes.c_void_p, ctypes.c_int, ctypes.c_float


def _lib():
    from reconnet_amd._lib import lib

    h = lib()
    h.rn_tv_sqs.restype = ci
    h.rn_tv_sqs.argtypes = [P, P, P, ci, ci, ci, cf, P, P]
    h.rn_tv_sqs_noalloc.restype = ci
    h.rn_tv_sqs_noalloc.argtypes = [P, P, P, ci, ci, ci, cf, P, P]
    return h


def _run(h, x, eps):
    import torch

    ns, nx, ny = x.shape
    t = torch.from_numpy(x).cuda()
    st = P(torch.cuda.current_stream().cuda_stream)
    g0, n0 = torch.empty_like(t), torch.empty_like(t)
    tv0 = (ctypes.c_double * ns)()
    rc = h.rn_tv_sqs(P(t.data_ptr()), P(g0.data_ptr()), P(n0.data_ptr()), nx, ny, ns, eps, ctypes.cast(tv0, P), st)
    assert rc == 0
    g1, n1 = torch.full_like(t, np.nan), torch.full_like(t, np.nan)
    tv1 = torch.full((ns,), np.nan, dtype=torch.float64, device="cuda")
    rc = h.rn_tv_sqs_noalloc(P(t.data_ptr()), P(g1.data_ptr()), P(n1.data_ptr()), nx, ny, ns, eps,
                             P(tv1.data_ptr()), st)
    assert rc == 0
    torch.cuda.synchronize()
    return (g0.cpu().numpy(), n0.cpu().numpy(), np.frombuffer(tv0, np.float64).copy(),
            g1.cpu().numpy(), n1.cpu().numpy(), tv1.cpu().numpy())


def test_noalloc_tv_equals_allocating_entry_point():
    import oracle as orc

    h = _lib()
    rng = np.random.default_rng(31)
    big = rng.uniform(0, 1, (5, 300, 257)).astype(np.float32)
    small = rng.uniform(0, 1, (2, 64, 50)).astype(np.float32)
    for x in (big, small):
        g0, n0, tv0, g1, n1, tv1 = _run(h, x, 1e-6)
        assert np.array_equal(g0, g1) and np.array_equal(n0, n1)
        assert np.array_equal(tv0, tv1)
        # and the values are the TV surrogate of the reference kernel
        go, no, tvo = orc.tv_sqs(x[-1])
        assert np.abs(g1[-1] - go).max() <= 1e-5 * np.abs(go).max()
        assert abs(tv1[-1] - tvo) <= 1e-9 * abs(tvo)


def test_noalloc_tv_rejects_sizes_beyond_its_array():
    import torch
    from reconnet_amd import last_error

    h = _lib()
    # 2 * ceil(1024/16)^2 * 33 = 270336 partials > 2^18
    nx = ny = 1024
    ns = 33
    t = torch.zeros((1,), device="cuda")
    st = P(torch.cuda.current_stream().cuda_stream)
    rc = h.rn_tv_sqs_noalloc(P(t.data_ptr()), P(t.data_ptr()), P(t.data_ptr()), nx, ny, ns, 1e-6, None, st)
    assert rc == -1
    assert "scratch array" in last_error()
    # the library stays usable
    x = np.random.default_rng(3).uniform(0, 1, (1, 32, 32)).astype(np.float32)
    g0, n0, tv0, g1, n1, tv1 = _run(h, x, 1e-6)
    assert np.array_equal(g0, g1) and np.array_equal(tv0, tv1)
