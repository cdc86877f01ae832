"""The allocation-free dot product (mgx_dot_noalloc) against the allocating one (mgx_dot).

Both launch the same two kernels with the same grid, so the results must agree bit for bit; the
allocation-free call keeps its partial sums in the library's per-GPU scratch array, which the
second, smaller call reuses (it must not read partials the first call left behind).  Calls go
through the C ABI."""
import ctypes as C

import numpy as np
import pytest

from compressiontests_amd import _lib as L
from compressiontests_amd.device import DeviceArray


def _dot(entry, a, b):
    res = DeviceArray(1)
    L.call(entry, a.n, a.vp, b.vp, res.vp, None)
    L.call("mgx_device_synchronize")
    return res.numpy()[0]


@pytest.mark.gpu
def test_dot_noalloc_matches_dot_on_two_calls(gpu):
    rng = np.random.default_rng(20261204)
    # 3,000,000 entries fill the capped grid (2048 partials); 1,000 entries use one block, so the
    # second call would pick up stale partials of the first if it read beyond its own
    for n in (3_000_000, 1_000):
        a_h, b_h = rng.standard_normal(n), rng.standard_normal(n)
        a, b = DeviceArray.from_numpy(a_h), DeviceArray.from_numpy(b_h)
        ref = _dot("mgx_dot", a, b)
        got = _dot("mgx_dot_noalloc", a, b)
        print(f"dot n={n}: mgx_dot {ref!r} mgx_dot_noalloc {got!r}", flush=True)
        assert got.tobytes() == ref.tobytes(), (n, got, ref)
        exact = float(np.dot(a_h.astype(np.longdouble), b_h.astype(np.longdouble)))
        assert abs(got - exact) <= 1e-12 * float(np.abs(a_h) @ np.abs(b_h))


@pytest.mark.gpu
def test_dot_noalloc_rejects_bad_arguments(gpu):
    res = DeviceArray(1)
    with pytest.raises(L.MgxInvalidArgument):
        L.call("mgx_dot_noalloc", -1, None, None, res.vp, None)
    with pytest.raises(L.MgxInvalidArgument):
        L.call("mgx_dot_noalloc", 4, None, None, res.vp, None)
    # n = 0: one block writes one zero partial
    z = DeviceArray(1)
    L.call("mgx_dot_noalloc", 0, z.vp, z.vp, res.vp, None)
    L.call("mgx_device_synchronize")
    assert res.numpy()[0] == 0.0
