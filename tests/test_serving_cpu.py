"""Host logic of the batched tracker and the tracking server (no device calls): the tile-class
partition, the pool's free list, the page-locked buffer sizes, the threaded host copies and the
stage-once hand-over between the staging thread and admission."""
import threading

import numpy as np
import pytest

from pypatools_amd import _lib
from pypatools_amd.batch import class_order
from pypatools_amd.serving import Job, _FreeList, _HostCopier, _PinnedPool, _stage_job


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


class _HostBuf:
    """Stand-in for a page-locked buffer (PinnedArray's view())."""

    def __init__(self, nbytes):
        self.nbytes = nbytes
        self.mem = np.zeros(nbytes, dtype=np.uint8)

    def view(self, dtype, count, offset_bytes=0):
        return self.mem[offset_bytes:].view(dtype)[:count]


class _Pool:
    def __init__(self):
        self.taken = 0
        self.lock = threading.Lock()

    def take(self, nbytes):
        with self.lock:
            self.taken += 1
        return _HostBuf(_PinnedPool._size(nbytes))


def _job(n, seed):
    rng = np.random.default_rng(seed)
    return Job(rng.normal(size=(n, 3)), rng.normal(size=(n, 3)), 1.0, "boris", 10, 1e-11)


def test_stage_once_then_hand_over():
    pool, cp = _Pool(), _HostCopier(2)
    j = _job(1000, 1)
    pin, copied = _stage_job(j, pool, cp)                  # the staging thread's copy
    assert copied and pool.taken == 1
    h = pin.view(np.float64, 6 * j.n)
    np.testing.assert_array_equal(h[:3 * j.n], j.r0.reshape(-1))
    np.testing.assert_array_equal(h[3 * j.n:], j.v0.reshape(-1))
    pin2, copied2 = _stage_job(j, pool, cp, take=True)     # admission takes the same buffer, no second copy
    assert pin2 is pin and not copied2 and pool.taken == 1
    assert _stage_job(j, pool, cp) == (None, False)       # a late staging call after the upload does nothing
    assert pool.taken == 1


def test_admission_first_copies_inline():
    pool, cp = _Pool(), _HostCopier(1)
    j = _job(10, 2)
    pin, copied = _stage_job(j, pool, cp, take=True)
    assert copied and pool.taken == 1
    np.testing.assert_array_equal(pin.view(np.float64, 3 * j.n), j.r0.reshape(-1))
    assert _stage_job(j, pool, cp) == (None, False)


def test_stager_and_admission_race_copy_once():
    """Many jobs staged by a thread while the main thread admits them in order: each is copied exactly
    once and every admitted buffer holds its own job's data."""
    pool, cp = _Pool(), _HostCopier(2)
    jobs = [_job(int(n), k) for k, n in enumerate(np.random.default_rng(9).integers(1, 40000, 60))]
    t = threading.Thread(target=lambda: [_stage_job(j, pool, cp) for j in jobs])
    t.start()
    copies = 0
    for j in jobs:
        pin, copied = _stage_job(j, pool, cp, take=True)
        copies += copied
        h = pin.view(np.float64, 6 * j.n)
        np.testing.assert_array_equal(h[:3 * j.n], j.r0.reshape(-1))
        np.testing.assert_array_equal(h[3 * j.n:], j.v0.reshape(-1))
    t.join()
    assert pool.taken == len(jobs)


def test_job_is_submitted_once():
    j = _job(4, 3)
    j.state = "queued"

    class _S:                                              # submit()'s checks only, without a device
        pool_n = 100
        maps = [None]
    from pypatools_amd.serving import TrackingServer
    with pytest.raises(ValueError, match="submitted once"):
        TrackingServer.submit(_S(), j)
