"""Pusher orchestration restated: push / push_batch / track / track_batch.  Test infrastructure only.

Follows src/PyPATools/pusher.py:
  * algorithm dispatch             push_batch :699-801, push :508-565
  * RK4 with 4 field evaluations   _rk4_step_batch :803-892, _rk4_step_single :567-618
  * track (single particle)        :620-693  (zeros history, Boris half steps)
  * track_batch                    :894-1012 (NaN history, active mask,
                                   Boris half step back/forward, record every n)

Two execution forms (see oracle/kernels.py): ``as_shipped=False`` (canonical,
vectorised, ``x*x``) and ``as_shipped=True`` (NumPy scalar ``pow`` squares and
the no-numba finalisation of track_batch, including its indexing of the
active-subset field arrays by full particle index, pusher.py:1000-1005).

Extensions (no reference symbol): algorithms 'boris_rel' and 'vay' (state
u = gamma v, staggered like 'boris'); time-dependent fields (objects with
``time_dependent = True`` called as ``f(pts, t)``), with t_n = t0 + n*dt and
RK4 stage times t_n + c*dt, c in (0, 0.5, 0.5, 1.0).
"""
import numpy as np

from . import kernels as K

REFERENCE_ALGORITHMS = ["leapfrog", "boris", "rk4", "yoshida", "vay_rel", "rk4_rel", "yoshida_rel"]
EXT_ALGORITHMS = ["boris_rel", "vay"]
ALGORITHMS = REFERENCE_ALGORITHMS + EXT_ALGORITHMS
STAGGERED = ("boris", "boris_rel", "vay")
U_STATE = ("boris_rel", "vay")


def _eval(field, pts, t):
    if getattr(field, "time_dependent", False):
        return field(pts, t)
    return field(pts)


class OraclePusher:
    def __init__(self, q_over_m, algorithm="boris", as_shipped=False):
        algorithm = algorithm.lower()
        if algorithm not in ALGORITHMS:
            raise ValueError(f"Unknown algorithm '{algorithm}'. Must be one of {ALGORITHMS}")
        self.q_over_m = q_over_m
        self.algorithm = algorithm
        self.relativistic = algorithm.endswith("_rel") or algorithm == "vay"
        self.as_shipped = as_shipped
        self.sq = K.sq_pow if as_shipped else K.sq_mul

    # -- one step ---------------------------------------------------------
    def _kick_u(self, u, E, B, dt):
        if self.algorithm == "boris_rel":
            return K.boris_rel_kick_ext(u, E, B, dt, self.q_over_m)
        return K.vay_kick_ext(u, E, B, dt, self.q_over_m)

    def _half_kick(self, v_or_u, E, B, dt):
        """Velocity-only kick used for the staggering half steps."""
        if self.algorithm == "boris":
            return K.boris_kick(v_or_u, E, B, dt, self.q_over_m, self.sq)
        return self._kick_u(v_or_u, E, B, dt)

    def step(self, r, v, efield, bfield, dt, t=0.0, batch=True):
        """push_batch (batch=True) / push (batch=False) on (M,3) arrays; returns new arrays."""
        a, qm, sq = self.algorithm, self.q_over_m, self.sq
        if a == "rk4" or a == "rk4_rel":
            def fe(pts, stage):
                ts = t + K.rk4_stage_times()[stage] * dt if np.ndim(dt) == 0 else t
                return _eval(efield, pts, ts), _eval(bfield, pts, ts)
            # batch NumPy derivative squares with array ``**2`` (exact); the
            # single-particle one goes through relativistic_gamma (scalar pow).
            sq_d = K.sq_mul if batch else sq
            return K.rk4_step(r, v, fe, dt, qm, a == "rk4_rel", sq_d)
        E = _eval(efield, r, t)
        B = _eval(bfield, r, t)
        if a == "leapfrog":
            return K.leapfrog_step(r, v, E, B, dt, qm)
        if a == "boris":
            return K.boris_step(r, v, E, B, dt, qm, sq)
        if a == "vay_rel":
            return K.vay_rel_step(r, v, E, B, dt, qm, sq)
        if a in ("yoshida", "yoshida_rel"):
            return K.yoshida_step(r, v, E, B, dt, qm, a == "yoshida_rel", sq)
        # extensions: v -> u -> kick -> v
        u_new = self._kick_u(K.u_from_v(v), E, B, dt)
        v_new = K.v_from_u(u_new)
        return r + v_new * dt, v_new

    def push_batch(self, r, v, efield, bfield, dt, t=0.0):
        return self.step(np.asarray(r, float), np.asarray(v, float), efield, bfield, dt, t, batch=True)

    def push(self, r, v, efield, bfield, dt, t=0.0):
        rn, vn = self.step(np.asarray(r, float).reshape(1, 3), np.asarray(v, float).reshape(1, 3),
                           efield, bfield, dt, t, batch=False)
        return rn[0], vn[0]

    # -- multi-step ---------------------------------------------------------
    def track_batch(self, r0, v0, efield, bfield, nsteps, dt, rec_every_n_steps=1,
                    t0=0.0, electrode_assembly=None):
        """pusher.py:894-1012 (+ extensions).  Returns (r_hist, v_hist, active)."""
        return self._track(r0, v0, efield, bfield, nsteps, dt, rec_every_n_steps, t0,
                           electrode_assembly, batch=True)

    def track(self, r0, v0, efield, bfield, nsteps, dt, rec_every_n_steps=1, t0=0.0):
        """pusher.py:620-693: single particle, zero-initialised history."""
        rh, vh, _ = self._track(np.asarray(r0, float).reshape(1, 3), np.asarray(v0, float).reshape(1, 3),
                                efield, bfield, nsteps, dt, rec_every_n_steps, t0, None, batch=False)
        return rh[:, 0, :], vh[:, 0, :]

    def _track(self, r0, v0, efield, bfield, nsteps, dt, rec, t0, elec, batch):
        a = self.algorithm
        r0 = np.asarray(r0, dtype=np.float64)
        v0 = np.asarray(v0, dtype=np.float64)
        M = r0.shape[0]
        n_records = nsteps // rec + 1
        fill = np.nan if batch else 0.0
        r_hist = np.full((n_records, M, 3), fill)
        v_hist = np.full((n_records, M, 3), fill)
        active = np.ones(M, dtype=bool)
        r = r0.copy()
        v = v0.copy()
        ustate = a in U_STATE
        u = K.u_from_v(v) if ustate else None
        if a in STAGGERED:
            E = _eval(efield, r, t0)
            B = _eval(bfield, r, t0)
            if ustate:
                u = self._half_kick(u, E, B, -0.5 * dt)
                v = K.v_from_u(u)
            else:
                v = self._half_kick(v, E, B, -0.5 * dt)
        r_hist[0] = r
        v_hist[0] = v
        rec_idx = 1
        for step in range(nsteps):
            t_n = t0 + step * dt
            r_old = r[active].copy()
            if ustate:
                E = _eval(efield, r[active], t_n)
                B = _eval(bfield, r[active], t_n)
                u[active] = self._kick_u(u[active], E, B, dt)
                v[active] = K.v_from_u(u[active])
                r[active] = r[active] + v[active] * dt
            else:
                r[active], v[active] = self.step(r[active], v[active], efield, bfield, dt, t_n, batch)
            if (step + 1) % rec == 0 and rec_idx < n_records:
                r_hist[rec_idx][active] = r[active]
                v_hist[rec_idx][active] = v[active]
                rec_idx += 1
            if elec is not None:
                hit = elec.segment_intersects_surface(r_old, r[active])
                old_idx = np.where(active)[0]
                active[old_idx[hit["hit_mask"]]] = False
                if not np.any(active):
                    break
        if a in STAGGERED:
            t_end = t0 + nsteps * dt
            E = _eval(efield, r[active], t_end)
            B = _eval(bfield, r[active], t_end)
            if ustate:
                u[active] = self._half_kick(u[active], E, B, 0.5 * dt)
                v[active] = K.v_from_u(u[active])
            elif self.as_shipped and batch:
                # no-numba finalisation, pusher.py:1000-1005: the field arrays hold
                # only active particles but are indexed by the full index i.
                for i in range(M):
                    if active[i]:
                        v[i] = K.boris_kick(v[i:i + 1], E[i:i + 1], B[i:i + 1], 0.5 * dt,
                                            self.q_over_m, self.sq)[0]
            else:
                v[active] = self._half_kick(v[active], E, B, 0.5 * dt)
            if batch:
                v_hist[-1][active] = v[active]
            else:
                v_hist[-1] = v
        return r_hist, v_hist, active
