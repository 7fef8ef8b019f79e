"""Many independent bunches in one launch (persistent batched push).

Every bunch has its own particles, species (q/m), algorithm, step count, dt,
E/B scale factors and field map (one of up to four resident 6-component maps).
All bunches live in one concatenated SoA state on the device; a launch
advances every bunch over a window of steps, one 256-particle tile of one
bunch per thread block, so per-bunch branches are uniform per block.  Between
windows the state is cell-sorted within each bunch (the sort key carries the
bunch index in its high bits, so bunches keep their slot ranges).  The window
loop can be captured once as a HIP graph and replayed: the launch reads its
step window from a device counter that the graph advances.

Each bunch follows Pusher.track_batch (src/PyPATools/pusher.py:894-1012) with
a ScaledField scale on E and B (field.py:661-690): the staggered algorithms
(boris, boris_rel, vay) take the half-step kick back before their first step
and the half-step kick forward after their last, inside the batched launches
(pp_bunch_flags).  Results are per-bunch final positions and velocities (v,
also for u-state algorithms) in each bunch's original particle order: the
last records r_hist[-1], v_hist[-1] of track_batch.
"""
import ctypes
import time

import numpy as np

from . import _lib
from .device import DeviceArray, Stream, synchronize
from .engine import DEFAULT_REFINE_BITS, STAGGERED, U_STATE, _fused_map, motion_axis
from .interpolators import DeviceFieldMap

BATCH_ALGORITHMS = ("boris", "vay", "vay_rel", "boris_rel", "rk4", "rk4_rel", "leapfrog")


class Bunch:
    """One bunch: initial (r0, v0) arrays (M, 3), q/m, algorithm, steps, dt, field scales, map index."""

    def __init__(self, r0, v0, q_over_m, algorithm="boris", nsteps=1000, dt=1e-11, escale=1.0, bscale=1.0,
                 map_index=0):
        if algorithm not in BATCH_ALGORITHMS:
            raise ValueError(f"batched pushes support {BATCH_ALGORITHMS}, not '{algorithm}'")
        self.r0 = np.ascontiguousarray(r0, dtype=np.float64).reshape(-1, 3)
        self.v0 = np.ascontiguousarray(v0, dtype=np.float64).reshape(-1, 3)
        if self.r0.shape != self.v0.shape:
            raise ValueError("r0 and v0 differ in shape")
        self.q_over_m = float(q_over_m)
        self.algorithm = algorithm
        self.nsteps = int(nsteps)
        if self.nsteps < 0 or self.nsteps >= 2 ** 31:
            raise ValueError("nsteps out of range")
        self.dt = float(dt)
        self.escale = float(escale)
        self.bscale = float(bscale)
        self.map_index = int(map_index)

    @property
    def n(self):
        return self.r0.shape[0]


def class_order(tb, ts, bunches, algorithm_of=lambda b: b.algorithm):
    """Tiles reordered RK4 bunches' first, then the single-gather ones (a stable partition), and the
    runs (pp_tile_class, first tile, count) of the classes present."""
    if len(tb) == 0:
        return tb, ts, []
    rk4 = np.array([algorithm_of(b) in _lib.RK4_ALGS for b in bunches], dtype=bool)[tb]
    order = np.concatenate([np.flatnonzero(rk4), np.flatnonzero(~rk4)])
    n_rk4 = int(rk4.sum())
    runs = []
    if n_rk4:
        runs.append((_lib.TILES_RK4, 0, n_rk4))
    if n_rk4 < len(tb):
        runs.append((_lib.TILES_SINGLE_GATHER, n_rk4, len(tb) - n_rk4))
    return np.ascontiguousarray(tb[order]), np.ascontiguousarray(ts[order]), runs


def make_map(efield, bfield, precision=None):
    """The fused 6-component device map of an (efield, bfield) pair on one uniform 3-D grid
    (for precision 'fp32_mixed' without the FP64 coefficient table)."""
    m = _fused_map(efield, bfield, fp64=precision != "fp32_mixed")
    if m is None:
        raise ValueError("E and B must be gridded 3-D Fields on the same grid")
    return m


class BatchedTracker:
    def __init__(self, maps, precision="fp32_mixed", block_size=256, steps_per_launch=500, sort_every=500,
                 use_graph=False):
        if precision not in ("fp32_mixed", "fp64_fast"):
            raise ValueError("batched pushes run in 'fp32_mixed' or 'fp64_fast'")
        if not 1 <= len(maps) <= _lib.MAX_BATCH_MAPS:
            raise ValueError(f"1 to {_lib.MAX_BATCH_MAPS} maps")
        self.maps = [m if isinstance(m, DeviceFieldMap) else make_map(*m, precision=precision) for m in maps]
        self.precision = precision
        self.block_size = int(block_size)
        self.steps_per_launch = int(steps_per_launch)
        self.sort_every = int(sort_every)
        self.use_graph = bool(use_graph)
        self.bunches = []
        self.timings = {}

    # ---- state ----------------------------------------------------------------
    def load(self, bunches):
        self.bunches = list(bunches)
        for b in self.bunches:
            if not 0 <= b.map_index < len(self.maps):
                raise ValueError(f"bunch map_index {b.map_index} out of range")
        counts = np.array([b.n for b in self.bunches], dtype=np.int64)
        offsets = np.concatenate([[0], np.cumsum(counts)[:-1]]).astype(np.int64)
        self.n = int(counts.sum())
        if self.n >= 2 ** 31:
            raise ValueError("at most 2^31 - 1 particles per batch")
        table = (_lib.PPBunch * len(self.bunches))()
        for k, b in enumerate(self.bunches):
            t = table[k]
            t.offset, t.count = int(offsets[k]), int(counts[k])
            t.algorithm = _lib.ALG_IDS[b.algorithm]
            t.nsteps = b.nsteps
            t.map_index = b.map_index
            t.flags = (_lib.BUNCH_KICK_START | _lib.BUNCH_KICK_END) if b.algorithm in STAGGERED else 0
            t.dt, t.q_over_m, t.escale, t.bscale = b.dt, b.q_over_m, b.escale, b.bscale
        self.offsets, self.counts = offsets, counts
        self.table_host = table
        self.table = DeviceArray((ctypes.sizeof(table),), np.uint8)
        _lib.call("pp_memcpy_h2d", self.table.ptr, ctypes.addressof(table), ctypes.sizeof(table), None)
        nt = ctypes.c_int64()
        _lib.call("pp_batch_tiles", ctypes.addressof(table), len(self.bunches), self.block_size, None, None,
                  ctypes.byref(nt))
        tb = np.empty(max(nt.value, 1), dtype=np.int32)
        ts = np.empty(max(nt.value, 1), dtype=np.int64)
        _lib.call("pp_batch_tiles", ctypes.addressof(table), len(self.bunches), self.block_size, tb.ctypes.data,
                  ts.ctypes.data, ctypes.byref(nt))
        self.ntiles = nt.value
        tb, ts, self.tile_runs = class_order(tb[:self.ntiles], ts[:self.ntiles], self.bunches)
        self.tile_bunch = DeviceArray.from_host(tb if self.ntiles else np.zeros(1, np.int32))
        self.tile_start = DeviceArray.from_host(ts if self.ntiles else np.zeros(1, np.int64))
        r = np.concatenate([b.r0 for b in self.bunches]) if self.bunches else np.zeros((0, 3))
        v = np.concatenate([b.v0 for b in self.bunches]) if self.bunches else np.zeros((0, 3))
        from .engine import ParticleState
        self.state = ParticleState.from_host(r, v)
        for k, b in enumerate(self.bunches):
            if b.algorithm in U_STATE and b.n:
                o = int(offsets[k])
                _lib.call("pp_v_to_u", self.state.vx.view_ptr(o), self.state.vy.view_ptr(o),
                          self.state.vz.view_ptr(o), b.n, None)
        group = np.repeat(np.arange(len(self.bunches), dtype=np.int32), counts)
        self.group = DeviceArray.from_host(group if self.n else np.zeros(1, np.int32))
        m0 = self.maps[0]
        ncell = int(np.prod([max(s - 1, 1) for s in m0.shape[:2]]) * (m0.shape[2] if m0.periodic_z
                                                                      else max(m0.shape[2] - 1, 1)))
        cell_bits = max(1, int(np.ceil(np.log2(max(ncell, 2)))))
        group_bits = max(1, int(np.ceil(np.log2(max(len(self.bunches), 2)))))
        self.refine_bits = max(0, min(DEFAULT_REFINE_BITS, 32 - cell_bits - group_bits))
        self.group_shift = cell_bits + self.refine_bits
        if self.group_shift + group_bits > 32:
            self.sort_every = 0     # no room in the key for the bunch index
        self.refine_axis = motion_axis(np.concatenate([b.v0[:1000] for b in self.bunches])) if self.bunches else 2
        self.state.ensure_ids()
        self.max_steps = max((b.nsteps for b in self.bunches), default=0)
        self.step = 0
        return self

    def _launch(self, step0, nsteps, stream, step0_dev=None):
        """One pp_push_batched per tile class present (RK4 tiles first: they are the longest)."""
        for cls, first, count in self.tile_runs:
            _lib.call("pp_push_batched", ctypes.byref(self._args(step0, nsteps, step0_dev, (cls, first, count))),
                      stream)

    def _args(self, step0, nsteps, step0_dev=None, run=None):
        st = self.state
        a = _lib.PPBatchArgs()
        a.precision = _lib.PRECISIONS[self.precision]
        a.block_size = self.block_size
        a.n = self.n
        a.x, a.y, a.z, a.vx, a.vy, a.vz = st.x.ptr, st.y.ptr, st.z.ptr, st.vx.ptr, st.vy.ptr, st.vz.ptr
        a.bunches = self.table.ptr
        a.nbunch = len(self.bunches)
        a.nmaps = len(self.maps)
        cls, first, count = run if run is not None else (_lib.TILES_ALL, 0, self.ntiles)
        a.tile_bunch, a.tile_start = self.tile_bunch.view_ptr(first), self.tile_start.view_ptr(first)
        a.ntiles, a.tile_class = count, cls
        a.step0, a.nsteps = int(step0), int(nsteps)
        a.step0_dev = step0_dev.ptr if step0_dev is not None else None
        for k, m in enumerate(self.maps):
            a.maps[k] = m.handle
        return a

    def _sort(self, stream=None):
        """Cell sort within each bunch, enqueued on `stream` (no host synchronisation)."""
        st = self.state
        pa = st._state_args()
        ws = st._workspace()
        _lib.call("pp_sort_state_ex", pa, self.maps[0].handle, self.group.ptr, self.group_shift, self.refine_axis,
                  self.refine_bits, st.ids.ptr, ws.ptr, ws.nbytes, stream)
        st.sorted = True

    # ---- running ---------------------------------------------------------------
    def run(self, nsteps=None, stream=None):
        """Advance every bunch by up to `nsteps` more steps (default: to the end)."""
        if self.n == 0:
            return
        end = self.max_steps if nsteps is None else min(self.max_steps, self.step + int(nsteps))
        spl = self.steps_per_launch
        t0 = time.perf_counter()
        if self.max_steps == 0 and self.step == 0:
            # only 0-step bunches: one window [0, 1) gives the staggered ones both half kicks
            self._launch(0, 1, stream)
            self.step = -1
        if self.use_graph:
            self._run_graph(end)
        else:
            since = None
            while self.step < end:
                if self.sort_every > 0 and (since is None or since >= self.sort_every):
                    self._sort(stream)
                    since = 0
                k = min(spl, end - self.step)
                if self.sort_every > 0:
                    k = min(k, self.sort_every - since)
                self._launch(self.step, k, stream)
                self.step += k
                if self.sort_every > 0:
                    since += k
        synchronize(stream)
        self.timings["run_s"] = self.timings.get("run_s", 0.0) + time.perf_counter() - t0

    def _run_graph(self, end):
        """Capture [launch window, advance counter] once per window length and replay it; the cell sorts
        between windows are enqueued eagerly on the same stream (the radix sort's library calls stay out
        of the graph), so the host never waits inside the loop."""
        spl = self.steps_per_launch
        if self.sort_every > 0:
            spl = min(spl, self.sort_every)
        s = Stream()
        synchronize()                   # work enqueued on the library stream (load, conversions) comes first
        counter = DeviceArray((1,), np.int64)
        counter.upload(np.array([self.step], dtype=np.int64))
        graphs = {}

        def graph_for(k):
            if k not in graphs:
                _lib.call("pp_graph_begin", s.handle)
                try:
                    self._launch(0, k, s.handle, counter)
                    _lib.call("pp_add_i64", counter.ptr, k, s.handle)
                finally:
                    g = ctypes.c_void_p()
                    _lib.call("pp_graph_end", s.handle, ctypes.byref(g))
                graphs[k] = g.value
            return graphs[k]

        since = None
        try:
            while self.step < end:
                if self.sort_every > 0 and (since is None or since >= self.sort_every):
                    self._sort(s.handle)        # eager, ordered on the replay stream: no host wait
                    since = 0
                k = min(spl, end - self.step)
                if self.sort_every > 0:
                    k = min(k, self.sort_every - since)
                _lib.call("pp_graph_launch", graph_for(k), s.handle)
                self.step += k
                if self.sort_every > 0:
                    since += k
            s.sync()
        finally:
            for g in graphs.values():
                _lib.lib().pp_graph_destroy(g)
        self.graph_count = len(graphs)

    def results(self):
        """Per-bunch (r, v) host arrays in original particle order (v = velocity for every algorithm)."""
        st = self.state
        st.unsort()
        for k, b in enumerate(self.bunches):
            if b.algorithm in U_STATE and b.n:
                o = int(self.offsets[k])
                _lib.call("pp_u_to_v", st.vx.view_ptr(o), st.vy.view_ptr(o), st.vz.view_ptr(o), b.n, None)
        r, v = st.to_host()
        for k, b in enumerate(self.bunches):
            if b.algorithm in U_STATE and b.n:
                o = int(self.offsets[k])
                _lib.call("pp_v_to_u", st.vx.view_ptr(o), st.vy.view_ptr(o), st.vz.view_ptr(o), b.n, None)
        out = []
        for k, b in enumerate(self.bunches):
            o, c = int(self.offsets[k]), int(self.counts[k])
            out.append((r[o:o + c], v[o:o + c]))
        return out

    @property
    def particle_steps(self):
        return int(sum(b.n * b.nsteps for b in self.bunches))


def c4_bunches(nbunch=512, seed=4, nmin=1e3, nmax=5e5, smin=1000, smax=10000, n=128, scale=1.0):
    """C4: log-uniform bunch sizes, per-bunch scale in [0.5, 1.5], species from
    {proton 2.5 MeV, H2_1+ 1 MeV, electron 100 keV} along +z, steps uniform in
    [smin, smax], Boris or Vay per bunch; dt per species so a particle crosses
    about one 1 mm cell per step.  `scale` shrinks the particle counts (tests)."""
    from .species import IonSpecies
    from .workloads import velocity_from_kinetic
    rng = np.random.default_rng(seed)
    species = [("proton", 2.5, 5e-11), ("H2_1+", 1.0, 1e-10), ("electron", 0.1, 6e-12)]
    zc = (n - 1) / 2.0 * 1e-3
    out = []
    for _ in range(nbunch):
        M = max(1, int(round(np.exp(rng.uniform(np.log(nmin), np.log(nmax))) * scale)))
        sp, ke, dt = species[rng.integers(3)]
        speed = velocity_from_kinetic(ke, sp)
        r0 = np.column_stack([rng.normal(0, 2e-3, M), rng.normal(0, 2e-3, M), rng.normal(zc, 2e-3, M)])
        vz = speed * (1 + 1e-3 * rng.standard_normal(M))
        v0 = np.column_stack([np.zeros(M), np.zeros(M), vz])
        alg = "boris" if rng.uniform() < 0.5 else "vay"
        steps = int(rng.integers(smin, smax + 1))
        s = float(rng.uniform(0.5, 1.5))
        out.append(Bunch(r0, v0, IonSpecies(sp).q_over_m, alg, steps, dt, s, s, 0))
    return out
