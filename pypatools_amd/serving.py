"""Continuous-batching tracking server on one GPU.

Jobs (particles + algorithm + steps + map) arrive at any time.  The server owns
one device-resident particle pool (SoA float64, sized from free HBM at start)
and a free list of slot ranges.  Each scheduling round it

  1. admits queued jobs into free pool ranges (FIFO, a job that does not fit
     yet waits while smaller ones behind it may go ahead, at most
     ``max_bypass`` times) -- admission is by free HBM: the pool is the HBM
     budget for particle state;
  2. advances every running job by one chunk (``chunk_steps``, 256 by
     default) in ONE batched launch, so the work of all running jobs overlaps
     on the device while their state stays resident between chunks;
  3. retires finished jobs: results to the host, slot range freed.

Each job's particles are cell-sorted inside its own range at admission (and
every ``sort_every_chunks`` chunks) and restored to the caller's order at
retirement, so a job's result does not depend on which other jobs shared its
chunks (tests/test_gpu_batch_serve.py checks alone vs in the mix, bit for bit).
"""
import collections
import ctypes
import threading
import time

import numpy as np

from . import _lib
from .batch import BATCH_ALGORITHMS, make_map
from .device import DeviceArray, mem_info, synchronize
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


class TrackingServer:
    def __init__(self, maps, precision="fp32_mixed", chunk_steps=256, pool_particles=None, hbm_fraction=0.5,
                 block_size=256, sort_every_chunks=4, max_bypass=8):
        if precision not in ("fp32_mixed", "fp64_fast"):
            raise ValueError("serving runs in 'fp32_mixed' or 'fp64_fast'")
        self.maps = [m if isinstance(m, DeviceFieldMap) else make_map(*m) for m in maps]
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
        self.finished = []
        self.lock = threading.Lock()
        self.rounds = 0
        self.particle_steps = 0
        self._ws = None
        self._ws_bytes = 0
        self._stage = None

    # ---- client side -------------------------------------------------------------
    def submit(self, job):
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
            return not self.queue and not self.running

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

    def _workspace(self, n):
        if self._ws_bytes < 1 or self._ws_need(n) > self._ws_bytes:
            b = self._ws_need(max(n, 1 << 20))
            self._ws = DeviceArray((b,), np.uint8)
            self._ws_bytes = b
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
        job.offset = offset
        p, o, n = self.pool, offset, job.n
        stage = DeviceArray.from_host(job.r0)
        _lib.call("pp_aos_to_soa", stage.ptr, n, p.x.view_ptr(o), p.y.view_ptr(o), p.z.view_ptr(o), None)
        stage.upload(job.v0)
        _lib.call("pp_aos_to_soa", stage.ptr, n, p.vx.view_ptr(o), p.vy.view_ptr(o), p.vz.view_ptr(o), None)
        _lib.call("pp_iota_i32", p.ids.view_ptr(o), n, None)
        if job.algorithm in U_STATE:
            _lib.call("pp_v_to_u", p.vx.view_ptr(o), p.vy.view_ptr(o), p.vz.view_ptr(o), n, None)
        if n >= 4096:
            self._sort_job(job)
        synchronize()           # the staging buffer is released on return
        job.state = "running"
        job.t_admit = time.perf_counter()

    def _retire(self, job):
        p, o, n = self.pool, job.offset, job.n
        ws = self._workspace(n)
        _lib.call("pp_unsort_state", self._range_args(job), p.ids.view_ptr(o), ws.ptr, ws.nbytes, None)
        if job.algorithm in U_STATE:
            _lib.call("pp_u_to_v", p.vx.view_ptr(o), p.vy.view_ptr(o), p.vz.view_ptr(o), n, None)
        out = DeviceArray((n, 3))
        _lib.call("pp_soa_to_aos", p.x.view_ptr(o), p.y.view_ptr(o), p.z.view_ptr(o), n, out.ptr, None)
        r = out.to_host()
        _lib.call("pp_soa_to_aos", p.vx.view_ptr(o), p.vy.view_ptr(o), p.vz.view_ptr(o), n, out.ptr, None)
        v = out.to_host()
        job.result = (r, v)
        job.t_done = time.perf_counter()
        job.state = "done"
        self.free.free(o, n)
        job.offset = None
        job.event.set()

    # ---- scheduling ------------------------------------------------------------------
    def _admit_queued(self, now=None):
        admitted = 0
        with self.lock:
            pending = list(self.queue)
        bypassed = 0
        for job in pending:
            if job.arrival is not None and now is not None and job.arrival > now:
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
        return admitted

    def round(self, now=None):
        """One scheduling round: admit, advance all running jobs one chunk, retire finished ones."""
        self._admit_queued(now)
        if not self.running:
            return 0
        table = (_lib.PPBunch * len(self.running))()
        work = 0
        for k, job in enumerate(self.running):
            t = table[k]
            t.offset, t.count = job.offset, job.n
            t.algorithm = _lib.ALG_IDS[job.algorithm]
            t.nsteps = min(self.chunk_steps, job.nsteps - job.done_steps)
            t.map_index = job.map_index
            if job.algorithm in STAGGERED:     # track_batch's half kicks in the first and last chunk
                t.flags = ((_lib.BUNCH_KICK_START if job.done_steps == 0 else 0)
                           | (_lib.BUNCH_KICK_END if job.done_steps + t.nsteps == job.nsteps else 0))
            t.dt, t.q_over_m, t.escale, t.bscale = job.dt, job.q_over_m, job.escale, job.bscale
            work += job.n * t.nsteps
        nt = ctypes.c_int64()
        _lib.call("pp_batch_tiles", ctypes.addressof(table), len(table), self.block_size, None, None,
                  ctypes.byref(nt))
        tb = np.empty(max(nt.value, 1), dtype=np.int32)
        ts = np.empty(max(nt.value, 1), dtype=np.int64)
        _lib.call("pp_batch_tiles", ctypes.addressof(table), len(table), self.block_size, tb.ctypes.data,
                  ts.ctypes.data, ctypes.byref(nt))
        dtab = DeviceArray((ctypes.sizeof(table),), np.uint8)
        _lib.call("pp_memcpy_h2d", dtab.ptr, ctypes.addressof(table), ctypes.sizeof(table), None)
        dtb, dts = DeviceArray.from_host(tb), DeviceArray.from_host(ts)
        a = _lib.PPBatchArgs()
        p = self.pool
        a.precision = _lib.PRECISIONS[self.precision]
        a.block_size = self.block_size
        a.n = self.pool_n
        a.x, a.y, a.z, a.vx, a.vy, a.vz = p.x.ptr, p.y.ptr, p.z.ptr, p.vx.ptr, p.vy.ptr, p.vz.ptr
        a.bunches, a.nbunch, a.nmaps = dtab.ptr, len(table), len(self.maps)
        a.tile_bunch, a.tile_start, a.ntiles = dtb.ptr, dts.ptr, nt.value
        a.step0, a.nsteps = 0, self.chunk_steps
        for k, m in enumerate(self.maps):
            a.maps[k] = m.handle
        _lib.call("pp_push_batched", ctypes.byref(a), None)
        synchronize()
        self.rounds += 1
        self.particle_steps += work
        still = []
        for job in self.running:
            job.done_steps += min(self.chunk_steps, job.nsteps - job.done_steps)
            job.chunks += 1
            if job.done_steps >= job.nsteps:
                self._retire(job)
                self.finished.append(job)
            else:
                if self.sort_every_chunks and job.chunks % self.sort_every_chunks == 0 and job.n >= 4096:
                    self._sort_job(job)
                still.append(job)
        self.running = still
        return work

    def run_until_idle(self, max_rounds=10 ** 9):
        r = 0
        while not self.idle and r < max_rounds:
            self.round(time.perf_counter())
            r += 1


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
