"""The gather-bandwidth probe used for the measured gather roofline in bench.py."""
import numpy as np
import pytest

import pypatools_amd as pa
from pypatools_amd import _lib
from pypatools_amd.batch import make_map
from pypatools_amd.device import DeviceArray
from pypatools_amd.workloads import c2_map

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _device():
    from pypatools_amd import device
    if device.device_count() == 0:
        pytest.skip("no HIP device visible")


def test_gather_probe_reads_the_table():
    grid, E, B = c2_map(16)
    m = make_map(pa.Field.from_arrays(grid, E), pa.Field.from_arrays(grid, B))
    n = 64 * 40
    out = DeviceArray((n,), np.float32)
    _lib.call("pp_gather_probe", m.handle, 0, 64, 5, n, out.ptr, None)
    a = out.to_host().reshape(-1, 64)
    assert np.all(np.isfinite(a)) and np.any(a != 0)
    assert np.all(a == a[:, :1])                       # mode 0: one cell per wavefront
    _lib.call("pp_gather_probe", m.handle, 1, 64, 5, n, out.ptr, None)
    b = out.to_host().reshape(-1, 64)
    assert np.all(np.isfinite(b)) and np.any(b != b[:, :1])
    with pytest.raises(ValueError):
        _lib.call("pp_gather_probe", m.handle, 0, 48, 5, n, out.ptr, None)   # not a power of two
    with pytest.raises(ValueError):
        _lib.call("pp_gather_probe", m.handle, 2, 64, 5, n, out.ptr, None)   # unknown mode
