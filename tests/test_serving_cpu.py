"""Host logic of the batched tracker and the tracking server (no device calls): the tile-class
partition, the pool's free list, the page-locked buffer sizes, the threaded host copies and the
submit-once rule."""
import numpy as np
import pytest

from pypatools_amd import _lib
from pypatools_amd.batch import class_order
from pypatools_amd.serving import Job, _FreeList, _HostCopier, _PinnedPool


class _B:
    def __init__(self, algorithm):
        self.algorithm = algorithm


def test_class_order_puts_rk4_tiles_first_stably():
    bunches = [_B("boris"), _B("rk4"), _B("vay"), _B("rk4_rel")]
    tb = np.array([0, 0, 1, 2, 3, 3, 1, 0], dtype=np.int32)
    ts = np.arange(len(tb), dtype=np.int64) * 256
    tb2, ts2, runs = class_order(tb, ts, bunches)
    # RK4-class tiles (bunches 1 and 3) first, in their original order, then the rest in theirs
    assert list(tb2) == [1, 3, 3, 1, 0, 0, 2, 0]
    assert list(ts2) == [512, 1024, 1280, 1536, 0, 256, 768, 1792]
    assert runs == [(_lib.TILES_RK4, 0, 4), (_lib.TILES_SINGLE_GATHER, 4, 4)]
    # every tile kept exactly once
    assert sorted(zip(tb2, ts2)) == sorted(zip(tb, ts))


def test_class_order_single_class_and_empty():
    bunches = [_B("boris"), _B("vay")]
    tb = np.array([1, 0, 1], dtype=np.int32)
    ts = np.array([0, 256, 512], dtype=np.int64)
    tb2, ts2, runs = class_order(tb, ts, bunches)
    assert list(tb2) == [1, 0, 1] and runs == [(_lib.TILES_SINGLE_GATHER, 0, 3)]
    tb2, ts2, runs = class_order(tb, ts, [_B("rk4"), _B("rk4")])
    assert runs == [(_lib.TILES_RK4, 0, 3)]
    tb2, ts2, runs = class_order(tb[:0], ts[:0], bunches)
    assert runs == [] and len(tb2) == 0


def test_free_list_first_fit_and_merge():
    fl = _FreeList(100)
    a = fl.alloc(30)
    b = fl.alloc(50)
    c = fl.alloc(20)
    assert (a, b, c) == (0, 30, 80) and fl.alloc(1) is None
    fl.free(b, 50)
    assert fl.alloc(60) is None            # 50 free, not contiguous with anything else
    fl.free(a, 30)
    assert fl.ranges == [(0, 80)]          # merged with its neighbour
    assert fl.alloc(80) == 0 and fl.free_slots == 0
    fl.free(0, 80)
    fl.free(c, 20)
    assert fl.ranges == [(0, 100)]


def test_pinned_sizes_are_powers_of_two_from_1_mib():
    assert _PinnedPool._size(1) == 1 << 20
    assert _PinnedPool._size(1 << 20) == 1 << 20
    assert _PinnedPool._size((1 << 20) + 1) == 1 << 21
    assert _PinnedPool._size(48 * 2_000_000) == 1 << 27


@pytest.mark.parametrize("threads", [1, 3, 8])
def test_host_copier_copies_every_slice(threads):
    rng = np.random.default_rng(threads)
    src = rng.normal(size=3 * 1_000_003)               # above the split size, not a multiple of anything
    dst = np.zeros_like(src)
    cp = _HostCopier(threads)
    cp(dst, src)
    np.testing.assert_array_equal(dst, src)
    small = rng.normal(size=17)
    out = np.zeros_like(small)
    cp(out, small)
    np.testing.assert_array_equal(out, small)
    cp.close()


def test_job_is_submitted_once():
    rng = np.random.default_rng(3)
    j = Job(rng.normal(size=(4, 3)), rng.normal(size=(4, 3)), 1.0, "boris", 10, 1e-11)
    j.state = "queued"

    class _S:                                              # submit()'s checks only, without a device
        pool_n = 100
        maps = [None]
    from pypatools_amd.serving import TrackingServer
    with pytest.raises(ValueError, match="submitted once"):
        TrackingServer.submit(_S(), j)
