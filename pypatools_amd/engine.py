"""Device-resident tracking engine behind Pusher.

The state of a track lives on the device as SoA float64 arrays for the whole
run.  With fields the kernels can read directly (constants, gridded maps, their
scaling and RF factors) the step loop is one fused kernel per chunk of steps:
each particle stays in registers for the chunk and HBM is touched at chunk
entry/exit and for history records.  Other fields (sums on different grids,
arbitrary Python callables) and electrode collision tests run one step per
launch with the fields evaluated between launches.

Bookkeeping follows Pusher.track_batch (src/PyPATools/pusher.py:894-1012):
NaN-filled history of nsteps // rec + 1 records, record 0 = initial state
(velocity half a step back for the staggered schemes), record k after step
k*rec, the last velocity record overwritten with the resynchronised velocity,
an ``active`` mask updated by the electrode test after each step.
"""
import ctypes

import numpy as np

from . import _lib
from .device import DeviceArray
from .field import Field, RFField, ScaledField, compile_term, evaluate_on_device, is_time_dependent
from .interpolators import DeviceFieldMap

REFERENCE_ALGORITHMS = ["leapfrog", "boris", "rk4", "yoshida", "vay_rel", "rk4_rel", "yoshida_rel"]
EXTENSION_ALGORITHMS = ["boris_rel", "vay"]
STAGGERED = ("boris", "boris_rel", "vay")
U_STATE = ("boris_rel", "vay")
RK4_FAMILY = ("rk4", "rk4_rel")

# particle-steps per fused launch (keeps single launches around a second or less)
DEFAULT_LAUNCH_WORK = 2 ** 30
# steps between cell sorts of the state when a field is a gridded map
DEFAULT_SORT_EVERY = 500
SORT_MIN_PARTICLES = 4096
# sub-cell bins (2^bits) along the direction of motion in the sort key
# cell sorts also order each cell's particles in 2^3 bins along the direction of motion: a wavefront
# then starts in fewer cells and keeps to fewer (C2: 1.3 distinct cells per wavefront one step after a
# sort instead of 2.2, 4.0 after 100 steps instead of 5.1; 1.93e11 particle-steps/s against 1.87e11)
DEFAULT_REFINE_BITS = 3


def motion_axis(v):
    """Axis of the largest mean |velocity| component of an (M, 3) array."""
    v = np.asarray(v, dtype=np.float64).reshape(-1, 3)
    return int(np.argmax(np.abs(v[: min(len(v), 100000)]).mean(axis=0))) if len(v) else 2


class ParticleState:
    """SoA device state of n particles: x, y, z and vx, vy, vz (u for u-state algorithms)."""

    def __init__(self, n):
        self.n = int(n)
        self.x, self.y, self.z = (DeviceArray((self.n,)) for _ in range(3))
        self.vx, self.vy, self.vz = (DeviceArray((self.n,)) for _ in range(3))
        self.alive = None

    @classmethod
    def from_host(cls, r, v):
        r = np.ascontiguousarray(r, dtype=np.float64).reshape(-1, 3)
        v = np.ascontiguousarray(v, dtype=np.float64).reshape(-1, 3)
        if r.shape != v.shape:
            raise ValueError(f"positions {r.shape} and velocities {v.shape} differ in shape")
        st = cls(r.shape[0])
        if st.n:
            tmp = DeviceArray.from_host(r)
            _lib.call("pp_aos_to_soa", tmp.ptr, st.n, st.x.ptr, st.y.ptr, st.z.ptr, None)
            tmp.upload(v)
            _lib.call("pp_aos_to_soa", tmp.ptr, st.n, st.vx.ptr, st.vy.ptr, st.vz.ptr, None)
        return st

    def positions_aos(self, out=None):
        out = out if out is not None else DeviceArray((self.n, 3))
        _lib.call("pp_soa_to_aos", self.x.ptr, self.y.ptr, self.z.ptr, self.n, out.ptr, None)
        return out

    def velocities_aos(self, out=None):
        out = out if out is not None else DeviceArray((self.n, 3))
        _lib.call("pp_soa_to_aos", self.vx.ptr, self.vy.ptr, self.vz.ptr, self.n, out.ptr, None)
        return out

    def to_host(self):
        return self.positions_aos().to_host(), self.velocities_aos().to_host()

    def load_aos(self, r_dev=None, v_dev=None):
        if r_dev is not None:
            _lib.call("pp_aos_to_soa", r_dev.ptr, self.n, self.x.ptr, self.y.ptr, self.z.ptr, None)
        if v_dev is not None:
            _lib.call("pp_aos_to_soa", v_dev.ptr, self.n, self.vx.ptr, self.vy.ptr, self.vz.ptr, None)

    # -- cell sort -------------------------------------------------------------------
    def ensure_ids(self):
        if getattr(self, "ids", None) is None:
            self.ids = DeviceArray((self.n,), np.int32)
            _lib.call("pp_iota_i32", self.ids.ptr, self.n, None)
            self.sorted = False
        return self.ids

    def _workspace(self):
        ws = getattr(self, "ws", None)
        if ws is None:
            b = ctypes.c_size_t()
            _lib.call("pp_sort_workspace_bytes", self.n, ctypes.byref(b))
            ws = self.ws = DeviceArray((max(int(b.value), 8),), np.uint8)
        return ws

    def _state_args(self, fp=None):
        a = _lib.PPPushArgs()
        a.n = self.n
        a.x, a.y, a.z = self.x.ptr, self.y.ptr, self.z.ptr
        a.vx, a.vy, a.vz = self.vx.ptr, self.vy.ptr, self.vz.ptr
        if fp is not None:
            a.efield, a.bfield = fp.tE, fp.tB
        return a

    def sort(self, fp, refine_axis=2, refine_bits=0):
        """Reorder by map cell (gather locality); a no-op when neither field is a map.

        refine_bits > 0 also orders particles inside a cell in 2^bits bins along
        refine_axis (the direction of motion), so a wavefront stays in one cell
        for more steps and reads the map through the scalar cache.
        """
        if self.n == 0:
            return
        ids = self.ensure_ids()
        ws = self._workspace()
        _lib.call("pp_sort_state_ex", self._state_args(fp), None, None, 0, int(refine_axis), int(refine_bits),
                  ids.ptr, ws.ptr, ws.nbytes, None)
        self.sorted = True

    def reset_order(self):
        """Declare the current contents to be in original order (after an external reload)."""
        if getattr(self, "ids", None) is not None:
            _lib.call("pp_iota_i32", self.ids.ptr, self.n, None)
        self.sorted = False

    def compact(self, fp, refine_axis=2, refine_bits=0):
        """Live slots first (ordered by map cell when a field is a map), dead slots after;
        the alive mask becomes 1...1 0...0.  Returns the live count."""
        if self.n == 0:
            return 0
        ids = self.ensure_ids()
        ws = self._workspace()
        a = self._state_args(fp)
        a.alive = self.alive.ptr
        n_live = ctypes.c_int64()
        _lib.call("pp_compact_state", a, None, int(refine_axis), int(refine_bits), ids.ptr, ctypes.byref(n_live),
                  ws.ptr, ws.nbytes, None)
        self.sorted = True
        return int(n_live.value)

    def unsort(self):
        """Back to the original order (the alive mask too)."""
        if self.n == 0 or not getattr(self, "sorted", False):
            return
        ws = self._workspace()
        a = self._state_args()
        if self.alive is not None:
            a.alive = self.alive.ptr
        _lib.call("pp_unsort_state", a, self.ids.ptr, ws.ptr, ws.nbytes, None)
        self.sorted = False

    def v_to_u(self):
        _lib.call("pp_v_to_u", self.vx.ptr, self.vy.ptr, self.vz.ptr, self.n, None)

    def u_to_v(self):
        _lib.call("pp_u_to_v", self.vx.ptr, self.vy.ptr, self.vz.ptr, self.n, None)


def _base_field(f):
    while isinstance(f, (ScaledField, RFField)):
        f = f.field
    return f if isinstance(f, Field) else None


_FUSED_CACHE = {}


def _fused_map(fe, fb, fp64=True):
    """One 6-component device map for E and B on the same grid (one cell lookup per gather).
    fp64=False: without the FP64 coefficient table (maps used by FP32-mixed launches only)."""
    be, bb = _base_field(fe), _base_field(fb)
    if be is None or bb is None or be is bb:
        return None
    se, sb = be._grid_source(), bb._grid_source()
    if se is None or sb is None or se[2] != sb[2] or len(se[0]) != len(sb[0]):
        return None
    if not all(np.array_equal(a, b) for a, b in zip(se[0], sb[0])):
        return None
    key = (id(be), id(bb), bool(fp64))
    hit = _FUSED_CACHE.get(key)
    if hit is not None and hit[0] is be and hit[1] is bb:
        return hit[2]
    m = DeviceFieldMap(se[0], se[1] + sb[1], 0.0, se[2], fp64=fp64)
    _FUSED_CACHE[key] = (be, bb, m)
    return m


class FieldPair:
    """E and B compiled for the push kernels."""

    def __init__(self, efield, bfield):
        self.efield, self.bfield = efield, bfield
        self.tE, self.tB = compile_term(efield), compile_term(bfield)
        self.direct = self.tE is not None and self.tB is not None
        self.time_dependent = is_time_dependent(efield) or is_time_dependent(bfield)
        self.map = None
        if self.direct and self.tE.kind == _lib.FIELD_MAP and self.tB.kind == _lib.FIELD_MAP:
            fm = _fused_map(efield, bfield)
            if fm is not None:
                self.map = fm
                self.tE.map = fm.handle
                self.tE.comp0 = 0
                self.tB.map = fm.handle
                self.tB.comp0 = 3



def build_rf_table(field, t0, dt, n_rows):
    """Host table [n_rows][4 stages][cos, sin] for RFField ``field`` at t_n + c dt, c = (0, 1/2, 1/2, 1).

    Evaluated element by element with the same expressions the pusher uses for a
    time-aware field (t_n = t0 + n*dt, stage time t_n + c*dt, ``omega*t + phase``).
    """
    tab = np.empty((n_rows, 4, 2))
    for n in range(n_rows):
        t_n = t0 + n * dt
        for s, c in enumerate((0.0, 0.5, 0.5, 1.0)):
            ts = t_n + c * dt
            arg = field.omega * ts + field.phase
            tab[n, s, 0] = np.cos(arg)
            tab[n, s, 1] = np.sin(arg)
    return tab


def _rf_source(f):
    while isinstance(f, ScaledField):
        f = f.field
    return f if isinstance(f, RFField) else None


class Tracker:
    """One track of n particles through (efield, bfield) on the device."""

    def __init__(self, algorithm, precision, q_over_m, block_size=0, launch_work=DEFAULT_LAUNCH_WORK,
                 sort_every=DEFAULT_SORT_EVERY, refine_bits=DEFAULT_REFINE_BITS, refine_axis=2):
        if algorithm not in _lib.ALG_IDS:
            raise ValueError(f"unknown algorithm '{algorithm}'")
        if precision not in _lib.PRECISIONS:
            raise ValueError(f"unknown precision '{precision}'")
        self.algorithm = algorithm
        self.precision = precision
        self.qm = float(q_over_m)
        self.block_size = int(block_size)
        self.launch_work = int(launch_work)
        self.sort_every = int(sort_every)   # steps between cell sorts (0: never) for map fields
        self.refine_bits = int(refine_bits)
        self.refine_axis = int(refine_axis)

    # -- kernel launch helpers ---------------------------------------------------
    def _args(self, st, fp, dt, nsteps, step0, tE=None, tB=None, rf=None, hist=None, kick=False, dt_pp=None):
        a = _lib.PPPushArgs()
        a.algorithm = _lib.ALG_IDS[self.algorithm]
        a.precision = _lib.PRECISIONS[self.precision]
        a.n = st.n
        a.x, a.y, a.z = st.x.ptr, st.y.ptr, st.z.ptr
        a.vx, a.vy, a.vz = st.vx.ptr, st.vy.ptr, st.vz.ptr
        a.alive = st.alive.ptr if st.alive is not None else None
        a.ids = st.ids.ptr if getattr(st, "sorted", False) else None
        a.dt_pp = dt_pp.ptr if dt_pp is not None else None
        a.dt = float(dt)
        a.q_over_m = self.qm
        a.nsteps = int(nsteps)
        a.step0 = int(step0)
        a.efield = tE if tE is not None else fp.tE
        a.bfield = tB if tB is not None else fp.tB
        if rf is not None:
            a.rf_table, a.rf_step0, a.rf_nsteps = rf[0].ptr, rf[1], rf[2]
        if hist is not None:
            a.r_hist, a.v_hist, a.rec_every, a.n_records = hist[0].ptr, hist[1].ptr, hist[2], hist[3]
        a.kick_only = 1 if kick else 0
        a.block_size = self.block_size
        return a

    def _launch(self, args, stream=None):
        _lib.call("pp_push", args, stream)

    def _rf(self, fp, t0, dt, nsteps):
        """Device rf table covering steps 0..nsteps (both fields share it: cos and sin columns)."""
        if not fp.time_dependent:
            return None
        srcs = {_rf_source(fp.efield), _rf_source(fp.bfield)} - {None}
        omegas = {(s.omega, s.phase) for s in srcs}
        if len(omegas) != 1:
            raise NotImplementedError("E and B must share one (omega, phase) in a fused push")
        src = next(iter(srcs))
        tab = build_rf_table(src, t0, dt, nsteps + 1)
        return (DeviceArray.from_host(tab), 0, nsteps + 1)

    # -- the track ------------------------------------------------------------------
    def track(self, r0, v0, efield, bfield, nsteps, dt, rec_every_n_steps=1, *, batch=True, t0=0.0,
              electrode_assembly=None, dt_pp=None, verbose=False, return_state=False):
        alg = self.algorithm
        nsteps = int(nsteps)
        rec = int(rec_every_n_steps)
        if rec < 1:
            raise ValueError("rec_every_n_steps must be >= 1")
        r0 = np.ascontiguousarray(r0, dtype=np.float64).reshape(-1, 3)
        v0 = np.ascontiguousarray(v0, dtype=np.float64).reshape(-1, 3)
        M = r0.shape[0]
        n_records = nsteps // rec + 1
        fill = np.nan if batch else 0.0
        if M == 0:
            empty = np.full((n_records, 0, 3), fill)
            return empty, empty.copy(), np.ones(0, dtype=bool)
        st = ParticleState.from_host(r0, v0)
        fp = FieldPair(efield, bfield)
        self.refine_axis = motion_axis(v0)
        elec = electrode_assembly
        active = np.ones(M, dtype=bool)
        if elec is not None:
            st.alive = DeviceArray.from_host(np.ones(M, dtype=np.uint8))
        dtp = DeviceArray.from_host(np.asarray(dt_pp, dtype=np.float64)) if dt_pp is not None else None
        rh = DeviceArray((n_records, M, 3))
        vh = DeviceArray((n_records, M, 3))
        rh.fill(fill)
        vh.fill(fill)
        hist = (rh, vh, rec, n_records)
        rf = self._rf(fp, t0, dt, nsteps) if fp.direct else None
        ustate = alg in U_STATE
        if ustate:
            st.v_to_u()
        if alg in STAGGERED:
            self._kick(st, fp, -0.5 * dt, 0, rf, t0, dt)
        # record 0
        st.positions_aos(_row(rh, 0, M))
        self._velocity_record(st, _row(vh, 0, M), ustate)
        if fp.direct and elec is None:
            self._fused_loop(st, fp, nsteps, dt, 0, rf, hist, dtp, verbose)
            st.unsort()
        else:
            self._slow_loop(st, fp, nsteps, dt, t0, rf, hist, elec, active, dtp, verbose)
        if alg in STAGGERED:
            self._kick(st, fp, 0.5 * dt, nsteps, rf, t0, dt)
            vfin = DeviceArray((M, 3))
            self._velocity_record(st, vfin, ustate)
        r_hist = rh.to_host()
        v_hist = vh.to_host()
        if alg in STAGGERED:
            vf = vfin.to_host()
            if batch:
                v_hist[-1][active] = vf[active]
            else:
                v_hist[-1] = vf
        if verbose:
            print(f"Tracking complete: {nsteps} steps, {M} particles")
        if return_state:
            if ustate:
                st.u_to_v()
            return r_hist, v_hist, active, st
        return r_hist, v_hist, active

    def _uses_sort(self, st, fp, dtp):
        maps = fp.tE.kind == _lib.FIELD_MAP or fp.tB.kind == _lib.FIELD_MAP
        return self.sort_every > 0 and maps and st.n >= SORT_MIN_PARTICLES and dtp is None

    def _fused_loop(self, st, fp, nsteps, dt, step0, rf, hist, dtp, verbose=False, steps_per_launch=None,
                    timer=None):
        """Chunks of fused steps; cell sort at the start and every sort_every steps."""
        per = int(steps_per_launch or max(1, self.launch_work // max(st.n, 1)))
        sort = self._uses_sort(st, fp, dtp)
        if sort:
            per = min(per, self.sort_every)
        step, end = int(step0), int(step0) + int(nsteps)
        since = None
        while step < end:
            if sort and (since is None or since >= self.sort_every):
                st.sort(fp, self.refine_axis, self.refine_bits)
                since = 0
            k = min(per, end - step)
            if sort:
                k = min(k, self.sort_every - since)
            if verbose:
                _progress(step, end)
            args = self._args(st, fp, dt, k, step, rf=rf, hist=hist, dt_pp=dtp)
            if timer is not None:
                timer(lambda: self._launch(args))
            else:
                self._launch(args)
            step += k
            if sort:
                since += k
        return step

    def run_resident(self, st, fp, nsteps, dt, step0=0, rf=None, steps_per_launch=None, timer=None):
        """Advance a device-resident state by nsteps (no history, no staggering).

        Chunks of ``steps_per_launch`` steps per fused launch; ``timer(launch)``
        wraps each launch (the benchmark records HIP events around it).
        """
        if not fp.direct:
            raise ValueError("run_resident needs fields the kernels read directly")
        return self._fused_loop(st, fp, nsteps, dt, step0, rf, None, None, False, steps_per_launch, timer)

    def single_step(self, r, v, efield, bfield, dt, t=0.0, dt_pp=None):
        """push_batch (pusher.py:699-801): one step, no staggering, no history; returns host (r, v)."""
        r = np.ascontiguousarray(r, dtype=np.float64).reshape(-1, 3)
        v = np.ascontiguousarray(v, dtype=np.float64).reshape(-1, 3)
        if r.shape[0] == 0:
            return r.copy(), v.copy()
        st = ParticleState.from_host(r, v)
        fp = FieldPair(efield, bfield)
        dtp = DeviceArray.from_host(np.asarray(dt_pp, dtype=np.float64)) if dt_pp is not None else None
        ustate = self.algorithm in U_STATE
        if ustate:
            st.v_to_u()
        if self.algorithm in RK4_FAMILY and not fp.direct:
            self._rk4_stage_loop(st, fp, 1, dt, t, None, dtp, False)
        elif fp.direct:
            rf = self._rf(fp, t, dt, 1)
            self._launch(self._args(st, fp, dt, 1, 0, rf=rf, dt_pp=dtp))
        else:
            tE, tB = self._array_terms(st, fp, t)
            self._launch(self._args(st, fp, dt, 1, 0, tE=tE, tB=tB, dt_pp=dtp))
        if ustate:
            st.u_to_v()
        return st.to_host()

    def _velocity_record(self, st, out_aos, ustate):
        if not ustate:
            return st.velocities_aos(out_aos)
        tmp = [DeviceArray((st.n,)) for _ in range(3)]
        for d, s in zip(tmp, (st.vx, st.vy, st.vz)):
            d.copy_from(s)
        _lib.call("pp_u_to_v", tmp[0].ptr, tmp[1].ptr, tmp[2].ptr, st.n, None)
        _lib.call("pp_soa_to_aos", tmp[0].ptr, tmp[1].ptr, tmp[2].ptr, st.n, out_aos.ptr, None)
        return out_aos

    def _kick(self, st, fp, dt_kick, step, rf, t0, dt):
        if fp.direct:
            self._launch(self._args(st, fp, dt_kick, 1, step, rf=rf, kick=True))
            return
        t = t0 + step * dt
        tE, tB = self._array_terms(st, fp, t)
        self._launch(self._args(st, fp, dt_kick, 1, step, tE=tE, tB=tB, kick=True))

    def _array_terms(self, st, fp, t, pts=None):
        pts = pts if pts is not None else st.positions_aos()
        E = evaluate_on_device(fp.efield, pts, st.n, self.precision, t)
        B = evaluate_on_device(fp.bfield, pts, st.n, self.precision, t)
        tE, tB = _lib.PPField(), _lib.PPField()
        for term, arr in ((tE, E), (tB, B)):
            term.kind = _lib.FIELD_ARRAY
            term.array = arr.ptr
            term.scale1 = 1.0
        self._keep = (E, B)
        return tE, tB

    def _slow_loop(self, st, fp, nsteps, dt, t0, rf, hist, elec, active, dtp, verbose):
        alg = self.algorithm
        M = st.n
        if alg in RK4_FAMILY and not fp.direct:
            if elec is not None:
                raise NotImplementedError("electrode collisions with RK4 need fields the kernels read directly")
            return self._rk4_stage_loop(st, fp, nsteps, dt, t0, hist, dtp, verbose)
        from .electrodes import TriangleElectrodes
        if isinstance(elec, TriangleElectrodes) and fp.direct:
            return self._lossy_loop(st, fp, nsteps, dt, rf, hist, elec, active, dtp, verbose)
        r_old_host = None
        for step in range(nsteps):
            if verbose:
                _progress(step, nsteps)
            if elec is not None:
                r_old_host = st.positions_aos().to_host()
            if fp.direct:
                self._launch(self._args(st, fp, dt, 1, step, rf=rf, hist=hist, dt_pp=dtp))
            else:
                tE, tB = self._array_terms(st, fp, t0 + step * dt)
                self._launch(self._args(st, fp, dt, 1, step, tE=tE, tB=tB, hist=hist, dt_pp=dtp))
            if elec is not None:
                r_new = st.positions_aos().to_host()
                idx = np.where(active)[0]
                hit = elec.segment_intersects_surface(r_old_host[idx], r_new[idx])
                active[idx[np.asarray(hit["hit_mask"], dtype=bool)]] = False
                st.alive.upload(active.astype(np.uint8))
                if not np.any(active):
                    break

    def _lossy_loop(self, st, fp, nsteps, dt, rf, hist, elec, active, dtp, verbose):
        """track_batch with an electrode assembly (pusher.py:966-990) on the device.

        Fused launches test every step's segment against the mesh inside the step
        loop (pp_push_lossy).  Every sort_every steps the state is compacted: live
        slots first, ordered by map cell, dead slots after (pp_compact_state), and
        the following launches cover the live prefix only.  Particles never
        interact, so the order of slots changes no result; history rows follow ids.
        """
        M = st.n
        compact = dtp is None and self.sort_every > 0
        if compact:
            st.ensure_ids()
        n_live, step, since = M, 0, None
        while step < nsteps and n_live > 0:
            if compact and (since is None or since >= self.sort_every):
                n_live = st.compact(fp, self.refine_axis, self.refine_bits)
                since = 0
                if n_live == 0:
                    break
            k = min(max(1, self.launch_work // n_live), nsteps - step)
            if compact:
                k = min(k, self.sort_every - since)
                since += k
            if verbose:
                _progress(step, nsteps)
            a = self._args(st, fp, dt, k, step, rf=rf, hist=hist, dt_pp=dtp)
            a.n, a.hist_rows = n_live, M
            _lib.call("pp_push_lossy", a, elec.handle, None)
            step += k
        st.unsort()
        active[:] = st.alive.to_host().astype(bool)

    def _rk4_stage_loop(self, st, fp, nsteps, dt, t0, hist, dtp, verbose):
        """_rk4_step_batch (pusher.py:803-892) with fields evaluated between device stage kernels."""
        M = st.n
        rel = self.algorithm == "rk4_rel"
        rh, vh, rec, n_records = hist if hist is not None else (None, None, 1, 0)
        R = st.positions_aos()
        V = st.velocities_aos()
        if dtp is not None:
            dtv = dtp.to_host()
            hdt_pp = DeviceArray.from_host(0.5 * dtv)
            dt_pp_ = DeviceArray.from_host(dtv)
            w_pp = DeviceArray.from_host(dtv / 6.0)
        else:
            hdt_pp = dt_pp_ = w_pp = None
        hdt, w = 0.5 * dt, dt / 6.0
        bufs = {k: DeviceArray((M, 3)) for k in ("k1", "k2", "k3", "k4", "R2", "V2", "R3", "V3", "R4", "V4", "Rn", "Vn")}

        def deriv(Vs, Rs, tt, out):
            E = evaluate_on_device(fp.efield, Rs, M, self.precision, tt)
            B = evaluate_on_device(fp.bfield, Rs, M, self.precision, tt)
            if rel:
                _lib.call("pp_rk4_rel_dbetagamma_dt_batch", Vs.ptr, E.ptr, B.ptr, M, self.qm, out.ptr, None)
            else:
                _lib.call("pp_lorentz_batch", Vs.ptr, E.ptr, B.ptr, M, self.qm, out.ptr, None)

        def axpy(a, s, s_pp, b, out):
            _lib.call("pp_axpy_batch", a.ptr, s, s_pp.ptr if s_pp is not None else None, b.ptr, M, out.ptr, None)

        b = bufs
        for step in range(nsteps):
            if verbose:
                _progress(step, nsteps)
            t_n = t0 + step * dt
            deriv(V, R, t_n, b["k1"])
            axpy(R, hdt, hdt_pp, V, b["R2"])
            axpy(V, hdt, hdt_pp, b["k1"], b["V2"])
            deriv(b["V2"], b["R2"], t_n + 0.5 * dt, b["k2"])
            axpy(R, hdt, hdt_pp, b["V2"], b["R3"])
            axpy(V, hdt, hdt_pp, b["k2"], b["V3"])
            deriv(b["V3"], b["R3"], t_n + 0.5 * dt, b["k3"])
            axpy(R, dt, dt_pp_, b["V3"], b["R4"])
            axpy(V, dt, dt_pp_, b["k3"], b["V4"])
            deriv(b["V4"], b["R4"], t_n + 1.0 * dt, b["k4"])
            wp = w_pp.ptr if w_pp is not None else None
            _lib.call("pp_rk4_combine_batch", R.ptr, w, wp, V.ptr, b["V2"].ptr, b["V3"].ptr, b["V4"].ptr, M,
                      b["Rn"].ptr, None)
            _lib.call("pp_rk4_combine_batch", V.ptr, w, wp, b["k1"].ptr, b["k2"].ptr, b["k3"].ptr, b["k4"].ptr, M,
                      b["Vn"].ptr, None)
            if rel:
                _lib.call("pp_clamp_speed_batch", b["Vn"].ptr, M, b["Vn"].ptr, None)
            R, b["Rn"] = b["Rn"], R
            V, b["Vn"] = b["Vn"], V
            if (step + 1) % rec == 0 and (step + 1) // rec < n_records:
                k = (step + 1) // rec
                _row(rh, k, M).copy_from(R)
                _row(vh, k, M).copy_from(V)
        st.load_aos(R, V)


class _RowView:
    """A (M, 3) window of a (n_records, M, 3) device array (no ownership)."""

    def __init__(self, base, k, M):
        self.ptr = base.ptr + k * M * 3 * 8
        self.shape = (M, 3)
        self.nbytes = M * 3 * 8

    def copy_from(self, other):
        _lib.call("pp_memcpy_d2d", self.ptr, other.ptr, self.nbytes, None)


def _row(arr, k, M):
    return _RowView(arr, k, M)


def _progress(step, nsteps):
    if nsteps >= 10 and step % (nsteps // 10) == 0:
        print(f"Step {step}/{nsteps} ({100 * step / nsteps:.0f}%)")
