"""Continuous-batching tracking server on one GPU.

Jobs (particles + algorithm + steps + map) arrive at any time.  The server owns
one device-resident particle pool (SoA float64, sized from free HBM at start)
and a free list of slot ranges.  Each scheduling round it

  1. admits queued jobs into free pool ranges (FIFO, a job that does not fit
     yet waits while smaller ones behind it may go ahead, at most
     ``max_bypass`` times) -- admission is by free HBM: the pool is the HBM
     budget for particle state;
  2. advances every running job by one chunk (``chunk_steps``, 128 by
     default) in one batched launch per tile class (RK4 jobs, the others), so
     the work of all running jobs overlaps on the device while their state
     stays resident between chunks;
  3. retires finished jobs: results to the host, slot range freed.

Each job's particles are cell-sorted inside its own range at admission (and
every ``sort_every_chunks`` chunks) and restored to the caller's order at
retirement, so a job's result does not depend on which other jobs shared its
chunks (tests/test_gpu_batch_serve.py checks alone vs in the mix, bit for bit).
"""
import collections
import concurrent.futures
import ctypes
import os
import threading
import time

import numpy as np

from . import _lib
from .batch import BATCH_ALGORITHMS, class_order, make_map
from .device import DeviceArray, Event, PinnedArray, copy_d2h_async, copy_h2d_async, mem_info, synchronize
from .engine import DEFAULT_REFINE_BITS, STAGGERED, U_STATE, ParticleState, motion_axis
from .interpolators import DeviceFieldMap


class Job:
    """A tracking request; fills in result, timestamps and state as it runs."""

    _next_id = 0

    def __init__(self, r0, v0, q_over_m, algorithm="boris", nsteps=1000, dt=1e-11, map_index=0, escale=1.0,
                 bscale=1.0, arrival=None):
        if algorithm not in BATCH_ALGORITHMS:
            raise ValueError(f"serving supports {BATCH_ALGORITHMS}, not '{algorithm}'")
        self.r0 = np.ascontiguousarray(r0, dtype=np.float64).reshape(-1, 3)
        self.v0 = np.ascontiguousarray(v0, dtype=np.float64).reshape(-1, 3)
        if self.r0.shape != self.v0.shape:
            raise ValueError("r0 and v0 differ in shape")
        self.n = self.r0.shape[0]
        self.q_over_m = float(q_over_m)
        self.algorithm = algorithm
        self.nsteps = int(nsteps)
        self.dt = float(dt)
        self.map_index = int(map_index)
        self.escale, self.bscale = float(escale), float(bscale)
        self.arrival = arrival
        self.refine_axis = motion_axis(self.v0)
        self.id = Job._next_id
        Job._next_id += 1
        self.state = "new"
        self.offset = None
        self.done_steps = 0
        self.chunks = 0
        self.t_submit = self.t_admit = self.t_done = None
        self.result = None
        self.event = threading.Event()

    @property
    def latency(self):
        start = self.arrival if self.arrival is not None else self.t_submit
        return None if self.t_done is None else self.t_done - start

    def wait(self, timeout=None):
        self.event.wait(timeout)
        return self.result


class _FreeList:
    """First-fit free list of [offset, offset+count) ranges in the pool."""

    def __init__(self, size):
        self.ranges = [(0, size)]

    def alloc(self, count):
        for k, (o, c) in enumerate(self.ranges):
            if c >= count:
                if c == count:
                    del self.ranges[k]
                else:
                    self.ranges[k] = (o + count, c - count)
                return o
        return None

    def free(self, offset, count):
        self.ranges.append((offset, count))
        self.ranges.sort()
        merged = []
        for o, c in self.ranges:
            if merged and merged[-1][0] + merged[-1][1] == o:
                merged[-1] = (merged[-1][0], merged[-1][1] + c)
            else:
                merged.append((o, c))
        self.ranges = merged

    @property
    def free_slots(self):
        return sum(c for _, c in self.ranges)


class _PinnedPool:
    """Reusable page-locked host buffers (admission uploads, result downloads), in power-of-two sizes
    from 1 MiB so that buffers are reused across jobs of similar sizes (page-locking is slow)."""

    def __init__(self, prewarm=()):
        self.free = [PinnedArray(self._size(b)) for b in prewarm]

    @staticmethod
    def _size(nbytes):
        return 1 << max(20, (max(int(nbytes), 1) - 1).bit_length())

    def take(self, nbytes):
        best = None
        for k, b in enumerate(self.free):
            if b.nbytes >= nbytes and (best is None or b.nbytes < self.free[best].nbytes):
                best = k
        if best is not None:
            return self.free.pop(best)
        return PinnedArray(self._size(nbytes))

    def give(self, buf):
        self.free.append(buf)


class _HostCopier:
    """Host copies between job arrays and page-locked buffers, split over a few threads when large
    (NumPy releases the GIL while it copies, so the slices run in parallel)."""

    SPLIT_BYTES = 8 << 20

    def __init__(self, threads):
        self.threads = max(1, int(threads))
        self._pool = None

    def __call__(self, dst, src):
        if self.threads == 1 or dst.nbytes < self.SPLIT_BYTES:
            np.copyto(dst, src)
            return
        if self._pool is None:
            self._pool = concurrent.futures.ThreadPoolExecutor(self.threads, thread_name_prefix="pp-copy")
        k = min(self.threads, -(-dst.nbytes // self.SPLIT_BYTES))
        edges = np.linspace(0, len(dst), k + 1).astype(np.int64)
        futs = [self._pool.submit(np.copyto, dst[a:b], src[a:b]) for a, b in zip(edges[:-1], edges[1:])]
        for f in futs:
            f.result()

    def close(self):
        if self._pool is not None:
            self._pool.shutdown(wait=True)
            self._pool = None


class _RoundSlot:
    """Device and pinned tile tables of one round in flight, reusable once its event has completed."""

    def __init__(self):
        self.event = Event()
        self.pending = False
        self.table_pin = self.tiles_pin = None
        self.table_dev = self.tb_dev = self.ts_dev = None

    def ensure(self, nbunch, ntiles):
        tb_bytes = ctypes.sizeof(_lib.PPBunch) * max(nbunch, 1)
        tl_bytes = 12 * max(ntiles, 1)          # int32 bunch + int64 start per tile
        if self.table_pin is None or self.table_pin.nbytes < tb_bytes:
            self.table_pin = PinnedArray(2 * tb_bytes)
            self.table_dev = DeviceArray((2 * tb_bytes,), np.uint8)
        if self.tiles_pin is None or self.tiles_pin.nbytes < tl_bytes:
            n = 2 * max(ntiles, 1)             # even: the int64 starts begin 8-byte aligned at 4 n bytes
            self.tiles_pin = PinnedArray(12 * n)
            self.tiles_cap = n
            self.tb_dev = DeviceArray((n,), np.int32)
            self.ts_dev = DeviceArray((n,), np.int64)

    def tile_views(self, ntiles):
        """(tile_bunch int32, tile_start int64) views of the pinned tile table; starts at byte 4 * capacity."""
        return (self.tiles_pin.view(np.int32, ntiles),
                self.tiles_pin.view(np.int64, ntiles, offset_bytes=4 * self.tiles_cap))


class TrackingServer:
    """Continuous batching with asynchronous rounds.

    Every device operation is enqueued on the library's stream and the host does
    not wait for it: admissions upload through page-locked buffers, each round's
    tile tables go to per-round device buffers (``max_inflight`` rounds may be
    queued; the host blocks only to reuse the buffers of the round that many
    rounds back), and finished jobs are downloaded asynchronously and completed
    when an event recorded after their copies has passed.  Pool slots are
    freed at enqueue time: later work on the same stream runs after the
    retirement copies.

    A round admits jobs until their host data reaches ``admit_bytes_per_round``
    (at least one job when nothing is running), so a burst of arrivals is
    ingested over several rounds while the device already pushes the first
    ones, instead of holding the first launch back until every upload is staged.
    Host copies of large jobs into and out of page-locked memory are split over
    ``copy_threads`` threads.  (Copying every submitted job ahead of admission on
    a thread of its own measured slower: it page-locks a buffer per queued job
    while the round loop launches.)
    """

    def __init__(self, maps, precision="fp32_mixed", chunk_steps=128, pool_particles=None, hbm_fraction=0.5,
                 block_size=256, sort_every_chunks=4, max_bypass=8, max_inflight=2, prewarm_particles=(2 ** 20,),
                 admit_bytes_per_round=64 << 20, copy_threads=None):
        if precision not in ("fp32_mixed", "fp64_fast"):
            raise ValueError("serving runs in 'fp32_mixed' or 'fp64_fast'")
        self.maps = [m if isinstance(m, DeviceFieldMap) else make_map(*m, precision=precision) for m in maps]
        if not 1 <= len(self.maps) <= _lib.MAX_BATCH_MAPS:
            raise ValueError(f"1 to {_lib.MAX_BATCH_MAPS} maps")
        self.precision = precision
        self.chunk_steps = int(chunk_steps)
        self.block_size = int(block_size)
        self.sort_every_chunks = int(sort_every_chunks)
        self.max_bypass = int(max_bypass)
        if pool_particles is None:
            free, _total = mem_info()
            per = 6 * 8 + 4 + 24 + 48          # state, ids, sort scratch, staging
            pool_particles = int(free * hbm_fraction) // per
        self.pool_n = int(min(pool_particles, 2 ** 31 - 1))
        self.pool = ParticleState(self.pool_n)
        self.pool.ids = DeviceArray((self.pool_n,), np.int32)
        self.free = _FreeList(self.pool_n)
        self.queue = collections.deque()
        self.running = []
        self.retiring = []          # (job, event, pinned buffer) with downloads in flight
        self.finished = []
        self.lock = threading.Lock()
        self.rounds = 0
        self.particle_steps = 0
        self._ws = None
        self._ws_bytes = 0
        self._stage = None          # device AoS staging of admissions (2n x 3 doubles: r then v)
        self._out = None            # device AoS staging of retirements
        self._pinned = _PinnedPool([48 * n for n in prewarm_particles])   # r and v of jobs that large
        self.admit_bytes_per_round = int(admit_bytes_per_round)
        self._copy = _HostCopier(copy_threads if copy_threads is not None else min(8, os.cpu_count() or 1))
        self._uploads = []          # (event, pinned buffer) of admission uploads in flight
        self._slots = [_RoundSlot() for _ in range(max(1, int(max_inflight)))]

    # ---- client side -------------------------------------------------------------
    def submit(self, job):
        if job.state != "new":
            raise ValueError("a job is submitted once")
        if job.n > self.pool_n:
            raise ValueError(f"job of {job.n} particles exceeds the pool of {self.pool_n}")
        if not 0 <= job.map_index < len(self.maps):
            raise ValueError("map_index out of range")
        job.t_submit = time.perf_counter()
        job.state = "queued"
        with self.lock:
            self.queue.append(job)
        return job

    @property
    def idle(self):
        with self.lock:
            return not self.queue and not self.running and not self.retiring

    # ---- device helpers ------------------------------------------------------------
    def _range_args(self, job):
        o, p = job.offset, self.pool
        a = _lib.PPPushArgs()
        a.n = job.n
        a.x, a.y, a.z = p.x.view_ptr(o), p.y.view_ptr(o), p.z.view_ptr(o)
        a.vx, a.vy, a.vz = p.vx.view_ptr(o), p.vy.view_ptr(o), p.vz.view_ptr(o)
        t = _lib.PPField()
        t.kind = _lib.FIELD_MAP
        t.map = self.maps[job.map_index].handle
        t.scale1 = 1.0
        a.efield = t
        return a

    def _grow(self, attr, nbytes):
        """A device buffer of at least nbytes; growing waits for the stream (the old one may be in use)."""
        buf = getattr(self, attr)
        if buf is None or buf.nbytes < nbytes:
            synchronize()
            setattr(self, attr, DeviceArray((max(nbytes, 1 << 20),), np.uint8))
        return getattr(self, attr)

    def _workspace(self, n):
        need = self._ws_need(max(n, 1 << 20))
        if self._ws is None or self._ws.nbytes < need:
            self._ws = self._grow("_ws", need)
        return self._ws

    @staticmethod
    def _ws_need(n):
        b = ctypes.c_size_t()
        _lib.call("pp_sort_workspace_bytes", n, ctypes.byref(b))
        return int(b.value)

    def _sort_job(self, job):
        ws = self._workspace(job.n)
        _lib.call("pp_sort_state_ex", self._range_args(job), None, None, 0, job.refine_axis, DEFAULT_REFINE_BITS,
                  self.pool.ids.view_ptr(job.offset), ws.ptr, ws.nbytes, None)

    def _admit(self, job, offset):
        """Enqueue the job's upload into its pool range (r, v through page-locked memory), ids, u and sort."""
        job.offset = offset
        p, o, n = self.pool, offset, job.n
        nb = 48 * n
        pin = self._pinned.take(nb)
        h = pin.view(np.float64, 6 * n)
        self._copy(h[:3 * n], job.r0.reshape(-1))
        self._copy(h[3 * n:], job.v0.reshape(-1))
        stage = self._grow("_stage", nb)
        copy_h2d_async(stage.ptr, pin, nb)
        _lib.call("pp_aos_to_soa", stage.ptr, n, p.x.view_ptr(o), p.y.view_ptr(o), p.z.view_ptr(o), None)
        _lib.call("pp_aos_to_soa", stage.ptr + 24 * n, n, p.vx.view_ptr(o), p.vy.view_ptr(o), p.vz.view_ptr(o),
                  None)
        _lib.call("pp_iota_i32", p.ids.view_ptr(o), n, None)
        if job.algorithm in U_STATE:
            _lib.call("pp_v_to_u", p.vx.view_ptr(o), p.vy.view_ptr(o), p.vz.view_ptr(o), n, None)
        if n >= 4096:
            self._sort_job(job)
        ev = Event()
        ev.record()
        self._uploads.append((ev, pin))
        job.state = "running"
        job.t_admit = time.perf_counter()

    def _retire(self, job):
        """Enqueue the job's unsort and download; it completes when its event has passed (_poll)."""
        p, o, n = self.pool, job.offset, job.n
        ws = self._workspace(n)
        _lib.call("pp_unsort_state", self._range_args(job), p.ids.view_ptr(o), ws.ptr, ws.nbytes, None)
        if job.algorithm in U_STATE:
            _lib.call("pp_u_to_v", p.vx.view_ptr(o), p.vy.view_ptr(o), p.vz.view_ptr(o), n, None)
        nb = 48 * n
        out = self._grow("_out", nb)
        _lib.call("pp_soa_to_aos", p.x.view_ptr(o), p.y.view_ptr(o), p.z.view_ptr(o), n, out.ptr, None)
        _lib.call("pp_soa_to_aos", p.vx.view_ptr(o), p.vy.view_ptr(o), p.vz.view_ptr(o), n, out.ptr + 24 * n, None)
        pin = self._pinned.take(nb)
        copy_d2h_async(pin, out.ptr, nb)
        ev = Event()
        ev.record()
        self.retiring.append((job, ev, pin))
        job.state = "retiring"
        self.free.free(o, n)          # later work on the stream runs after these copies
        job.offset = None

    def _poll(self, block=False):
        """Complete jobs whose downloads have finished; release upload buffers.  block: wait for the oldest."""
        if block and self.retiring:
            self.retiring[0][1].sync()
        still = []
        for job, ev, pin in self.retiring:
            if ev.query():
                h = pin.view(np.float64, 6 * job.n)
                r, v = np.empty((job.n, 3)), np.empty((job.n, 3))
                self._copy(r.reshape(-1), h[:3 * job.n])
                self._copy(v.reshape(-1), h[3 * job.n:])
                job.result = (r, v)
                job.t_done = time.perf_counter()
                job.state = "done"
                self._pinned.give(pin)
                self.finished.append(job)
                job.event.set()
            else:
                still.append((job, ev, pin))
        self.retiring = still
        keep = []
        for ev, pin in self._uploads:
            if ev.query():
                self._pinned.give(pin)
            else:
                keep.append((ev, pin))
        self._uploads = keep

    # ---- scheduling ------------------------------------------------------------------
    def _admit_queued(self, now=None):
        admitted = 0
        with self.lock:
            pending = list(self.queue)
        bypassed = 0
        staged = 0
        for job in pending:
            if job.arrival is not None and now is not None and job.arrival > now:
                break
            if staged >= self.admit_bytes_per_round and (admitted or self.running):
                break
            off = self.free.alloc(job.n)
            if off is None:
                bypassed += 1
                if bypassed > self.max_bypass:
                    break
                continue
            with self.lock:
                self.queue.remove(job)
            self._admit(job, off)
            self.running.append(job)
            admitted += 1
            staged += 48 * job.n
        return admitted

    def round(self, now=None):
        """One scheduling round: admit, enqueue one chunk of every running job, enqueue the retirement of
        jobs that finish with it, and complete jobs whose downloads have arrived.  Returns the work enqueued."""
        self._admit_queued(now)
        if not self.running:
            self._poll(block=bool(self.retiring))
            return 0
        slot = self._slots[self.rounds % len(self._slots)]
        if slot.pending:
            slot.event.sync()           # its tables are about to be rewritten
            slot.pending = False
        nb = len(self.running)
        nt = ctypes.c_int64()
        slot.ensure(nb, 0)
        table = (_lib.PPBunch * nb).from_address(slot.table_pin.ptr)
        work = 0
        # tiles of the costliest jobs first (four gathers per RK4 step): the dispatcher hands out
        # workgroups in table order, so the launch does not end on a tail of long tiles
        order = sorted(self.running, key=lambda j: -(4 if j.algorithm in ("rk4", "rk4_rel") else 1)
                       * min(self.chunk_steps, j.nsteps - j.done_steps))
        for k, job in enumerate(order):
            t = table[k]
            t.offset, t.count = job.offset, job.n
            t.algorithm = _lib.ALG_IDS[job.algorithm]
            t.nsteps = min(self.chunk_steps, job.nsteps - job.done_steps)
            t.map_index = job.map_index
            t.flags = 0
            if job.algorithm in STAGGERED:     # track_batch's half kicks in the first and last chunk
                t.flags = ((_lib.BUNCH_KICK_START if job.done_steps == 0 else 0)
                           | (_lib.BUNCH_KICK_END if job.done_steps + t.nsteps == job.nsteps else 0))
            t.dt, t.q_over_m, t.escale, t.bscale = job.dt, job.q_over_m, job.escale, job.bscale
            work += job.n * t.nsteps
        _lib.call("pp_batch_tiles", slot.table_pin.ptr, nb, self.block_size, None, None, ctypes.byref(nt))
        ntiles = max(nt.value, 1)
        slot.ensure(nb, ntiles)
        table = (_lib.PPBunch * nb).from_address(slot.table_pin.ptr)
        tb, ts = slot.tile_views(ntiles)
        _lib.call("pp_batch_tiles", slot.table_pin.ptr, nb, self.block_size, tb.ctypes.data, ts.ctypes.data,
                  ctypes.byref(nt))
        # RK4 jobs' tiles first, each class to the kernel built for it (pp_tile_class)
        tb_o, ts_o, runs = class_order(tb[:nt.value], ts[:nt.value], order)
        tb[:nt.value] = tb_o
        ts[:nt.value] = ts_o
        copy_h2d_async(slot.table_dev.ptr, slot.table_pin, ctypes.sizeof(_lib.PPBunch) * nb)
        copy_h2d_async(slot.tb_dev.ptr, slot.tiles_pin, 4 * ntiles)
        copy_h2d_async(slot.ts_dev.ptr, slot.tiles_pin, 8 * ntiles, offset_bytes=4 * slot.tiles_cap)
        a = _lib.PPBatchArgs()
        p = self.pool
        a.precision = _lib.PRECISIONS[self.precision]
        a.block_size = self.block_size
        a.n = self.pool_n
        a.x, a.y, a.z, a.vx, a.vy, a.vz = p.x.ptr, p.y.ptr, p.z.ptr, p.vx.ptr, p.vy.ptr, p.vz.ptr
        a.bunches, a.nbunch, a.nmaps = slot.table_dev.ptr, nb, len(self.maps)
        a.step0, a.nsteps = 0, self.chunk_steps
        for k, m in enumerate(self.maps):
            a.maps[k] = m.handle
        for cls, first, count in runs:
            a.tile_bunch, a.tile_start = slot.tb_dev.view_ptr(first), slot.ts_dev.view_ptr(first)
            a.ntiles, a.tile_class = count, cls
            _lib.call("pp_push_batched", ctypes.byref(a), None)
        self.rounds += 1
        self.particle_steps += work
        still = []
        for job in self.running:
            job.done_steps += min(self.chunk_steps, job.nsteps - job.done_steps)
            job.chunks += 1
            if job.done_steps >= job.nsteps:
                self._retire(job)
            else:
                if self.sort_every_chunks and job.chunks % self.sort_every_chunks == 0 and job.n >= 4096:
                    self._sort_job(job)
                still.append(job)
        self.running = still
        slot.event.record()
        slot.pending = True
        self._poll()
        return work

    def run_until_idle(self, max_rounds=10 ** 9):
        r = 0
        while not self.idle and r < max_rounds:
            self.round(time.perf_counter())
            r += 1
        synchronize()
        self._poll()


def c5_jobs(n_jobs, rate, seed=5, nmin=1e3, nmax=2e6, smin=1e2, smax=2e4, t0=0.0, template=None):
    """C5 requests: Poisson arrivals at `rate` jobs/s, N log-uniform in [nmin, nmax], steps log-uniform in
    [smin, smax], pusher from {boris, vay, rk4}, one of 4 maps; particles are slices of a 2.5 MeV proton
    template bunch (C2 geometry), dt = 5e-11 s."""
    from .species import IonSpecies
    from .workloads import c2_bunch
    rng = np.random.default_rng(seed)
    if template is None:
        template = c2_bunch(M=int(nmax), seed=seed + 1)
    tr, tv = template
    qm = IonSpecies("proton").q_over_m
    t = t0
    jobs = []
    for _ in range(n_jobs):
        t += rng.exponential(1.0 / rate)
        N = int(round(np.exp(rng.uniform(np.log(nmin), np.log(nmax)))))
        steps = int(round(np.exp(rng.uniform(np.log(smin), np.log(smax)))))
        start = int(rng.integers(0, len(tr) - N + 1))
        alg = ("boris", "vay", "rk4")[rng.integers(3)]
        jobs.append(Job(tr[start:start + N], tv[start:start + N], qm, alg, steps, 5e-11, int(rng.integers(4)),
                        arrival=t))
    return jobs
