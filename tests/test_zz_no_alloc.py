"""pp_beam_moments_noalloc (fixed device scratch, no allocation) against pp_beam_moments.

Two calls in a row through the C ABI, the second on a smaller state, each compared bit for bit
with the workspace entry point on the same state: the scratch keeps the previous call's partials,
and a smaller call must not pick any of them up.
"""
import ctypes

import numpy as np
import pytest

from pypatools_amd import _lib

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _device():
    from pypatools_amd import device
    if device.device_count() == 0:
        pytest.skip("no HIP device visible")
    device.init(0)


def _state(n, seed):
    from pypatools_amd.engine import ParticleState
    rng = np.random.default_rng(seed)
    r = rng.normal(0.0, 2e-3, (n, 3)) + np.array([0.0, 0.0, 0.06])
    v = np.column_stack([rng.normal(0, 1e4, n), rng.normal(0, 1e4, n), 2.2e7 * (1 + 1e-3 * rng.normal(size=n))])
    return ParticleState.from_host(r, v)


def _moments(st, noalloc):
    from pypatools_amd.device import DeviceArray, synchronize
    sums, cen = DeviceArray((_lib.MOMENT_SUMS,)), DeviceArray((_lib.MOMENT_CENTRAL,))
    sums.fill(-1.0)
    cen.fill(-1.0)
    args = st._state_args()
    if noalloc:
        _lib.call("pp_beam_moments_noalloc", args, 0, 938.272, 0, sums.ptr, cen.ptr, None)
    else:
        nb = ctypes.c_size_t()
        _lib.call("pp_beam_workspace_bytes", ctypes.byref(nb))
        ws = DeviceArray((nb.value,), np.uint8)
        _lib.call("pp_beam_moments", args, 0, 938.272, 0, sums.ptr, cen.ptr, ws.ptr, ws.nbytes, None)
    synchronize()
    return sums.to_host(), cen.to_host()


def test_noalloc_moments_equal_workspace_moments_twice_shrinking():
    b = ctypes.c_size_t()
    _lib.call("pp_beam_noalloc_scratch_bytes", ctypes.byref(b))
    need = ctypes.c_size_t()
    _lib.call("pp_beam_workspace_bytes", ctypes.byref(need))
    assert b.value >= need.value
    for n, seed in ((1_000_003, 1), (70_001, 2), (1, 3)):
        st = _state(n, seed)
        s_ref, c_ref = _moments(st, noalloc=False)
        s_na, c_na = _moments(st, noalloc=True)
        np.testing.assert_array_equal(s_na, s_ref)
        np.testing.assert_array_equal(c_na, c_ref)
        assert s_na[0] == n


def test_noalloc_moments_reject_bad_arguments():
    from pypatools_amd.device import DeviceArray
    lib = _lib.lib()
    st = _state(16, 4)
    out = DeviceArray((_lib.MOMENT_SUMS,))
    assert lib.pp_beam_moments_noalloc(st._state_args(), 0, 938.272, 0, out.ptr, None, None) == _lib.PP_ERR_INVALID
    assert lib.pp_beam_moments_noalloc(st._state_args(), 0, -1.0, 0, out.ptr, out.ptr, None) == _lib.PP_ERR_INVALID
