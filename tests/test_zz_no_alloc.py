"""The allocation-free entry point mgldg_mean_zero_device_noalloc (static
device scratch) against mgldg_mean_zero_device (scratch grown on demand),
through the C ABI: bitwise equal on two calls in a row, the second smaller
than the first (the shared scratch keeps the first call's partials), and
the column-count limit rejected before any launch."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_noalloc_mean_zero_matches_allocating_entry(gctx):
    import torch
    from mgldg._lib import lib
    L = lib()
    rng = np.random.default_rng(17)
    for n, k in ((300_007, 7), (1_001, 3), (65_536, 1)):
        print(f"[no-alloc] mean_zero n={n} k={k}", flush=True)
        a = rng.standard_normal((n, k)) + 3.0 * rng.standard_normal(k)
        t_ref = torch.tensor(a, device="cuda")
        t_new = t_ref.clone()
        torch.cuda.synchronize()
        assert L.mgldg_mean_zero_device(gctx.h, n, ctypes.c_void_p(t_ref.data_ptr()), k) == 0
        assert L.mgldg_mean_zero_device_noalloc(gctx.h, n, ctypes.c_void_p(t_new.data_ptr()), k) == 0
        assert torch.equal(t_ref, t_new)
        got = t_new.cpu().numpy()
        assert np.abs(got - (a - a.mean(axis=0))).max() < 1e-12 * max(1.0, np.abs(a).max())


def test_noalloc_mean_zero_rejects_too_many_columns(gctx):
    import torch
    from mgldg._lib import lib
    t = torch.ones(16, 257, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    assert lib().mgldg_mean_zero_device_noalloc(gctx.h, 16, ctypes.c_void_p(t.data_ptr()), 257) == -1
    assert bool((t == 1).all())
