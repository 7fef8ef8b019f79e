"""ParticleDistribution (src/PyPATools/particles.py:123-843) with its collective
quantities computed on the device.

The distribution is held as a device SoA state: positions and u = gamma v =
c * (beta gamma).  Centroid, rms sizes and angles, correlations, energy mean and
spread, extremes and the 4-rms ellipse counts are GPU reductions
(``pp_beam_moments``, ``pp_beam_ellipse_counts``); emittances and Twiss
parameters follow from them on the host with the reference's formulas.

Momentum <-> velocity uses the vector form v = c p / sqrt(1 + |p|^2).  The
reference converts component by component, v_i = c p_i / sqrt(1 + p_i^2)
(particles.py:72-77, 393-402), which is exact only for momenta along one axis;
``to_pusher`` therefore hands the pusher physical velocities.
"""
import ctypes

import numpy as np

from . import _lib
from .constants import CLIGHT
from .device import DeviceArray
from .species import IonSpecies


class ParticleDistribution:
    """particles.py:123-268: x_vec (N, 3) [m] and p_vec (N, 3) [beta*gamma] of one species."""

    def __init__(self, species=None, x=None, y=None, z=None, px=None, py=None, pz=None, x_vec=None, p_vec=None,
                 q=0.0, f=0.0, recalculate=True, z_energy=False):
        self._species = species if species is not None else IonSpecies("proton")
        separate = any(a is not None for a in (x, y, z, px, py, pz))
        vector = x_vec is not None or p_vec is not None
        if separate and vector:
            raise ValueError("Cannot specify both separate coordinates (x,y,z,px,py,pz) "
                             "AND vector coordinates (x_vec, p_vec). Choose one method.")
        if vector:
            if x_vec is None or p_vec is None:
                raise ValueError("Both x_vec and p_vec must be provided together.")
            x_vec, p_vec = np.asarray(x_vec, dtype=np.float64), np.asarray(p_vec, dtype=np.float64)
            if x_vec.shape != p_vec.shape:
                raise ValueError(f"x_vec and p_vec must have same shape. Got x_vec: {x_vec.shape}, "
                                 f"p_vec: {p_vec.shape}")
            if x_vec.ndim != 2 or x_vec.shape[1] != 3:
                raise ValueError(f"x_vec and p_vec must have shape (N, 3). Got: {x_vec.shape}")
            xv, pv = x_vec.copy(), p_vec.copy()
        else:
            cols = [np.atleast_1d(np.asarray(a, dtype=np.float64)) if a is not None else np.zeros(1)
                    for a in (x, y, z, px, py, pz)]
            if len({len(a) for a in cols}) != 1:
                raise ValueError("All coordinate arrays must have same length. Got lengths: "
                                 + ", ".join(f"{k}={len(a)}" for k, a in zip(("x", "y", "z", "px", "py", "pz"), cols)))
            xv, pv = np.column_stack(cols[:3]), np.column_stack(cols[3:])
        self.q, self.f = q, f
        self.z_energy = bool(z_energy)
        self._host = (xv, pv)
        self._state = None
        self._u_state = True
        self.numpart = len(xv)
        self.alive_mask = np.ones(self.numpart, dtype=bool)
        self._m = None
        if recalculate:
            self.recalculate_all()

    # -- construction from / hand-off to the device engine ---------------------------------
    @classmethod
    def from_state(cls, state, species=None, u_state=False, z_energy=False):
        """A distribution over a device-resident ParticleState (no copy); u_state: slots hold gamma v."""
        d = cls.__new__(cls)
        d._species = species if species is not None else IonSpecies("proton")
        d.q = d.f = 0.0
        d.z_energy = bool(z_energy)
        d._host = None
        d._state = state
        d._u_state = bool(u_state)
        d.numpart = state.n
        d.alive_mask = np.ones(d.numpart, dtype=bool)
        d._m = None
        return d

    @classmethod
    def from_pusher(cls, r, v, species=None, z_energy=False):
        """From pusher arrays: positions r and physical velocities v, (N, 3)."""
        r, v = np.asarray(r, dtype=np.float64), np.asarray(v, dtype=np.float64)
        b2 = np.sum(v * v, axis=1) / CLIGHT ** 2
        p = v / CLIGHT / np.sqrt(1.0 - b2)[:, None]
        return cls(species, x_vec=r, p_vec=p, z_energy=z_energy)

    def to_pusher(self):
        """(r, v) for Pusher.track_batch: positions and physical velocities v = c p / sqrt(1 + |p|^2)."""
        x, p = self.x_vec, self.p_vec
        return x.copy(), CLIGHT * p / np.sqrt(1.0 + np.sum(p * p, axis=1))[:, None]

    def _device(self):
        if self._state is None:
            from .engine import ParticleState
            xv, pv = self._host
            self._state = ParticleState.from_host(xv, pv * CLIGHT)     # u = gamma v = c (beta gamma)
            self._u_state = True
        return self._state

    def _args(self):
        st = self._device()
        a = st._state_args()
        if not np.all(self.alive_mask):
            self._alive_dev = DeviceArray.from_host(self.alive_mask.astype(np.uint8))
            a.alive = self._alive_dev.ptr
        return a

    def _workspace(self):
        ws = getattr(self, "_ws", None)
        if ws is None:
            b = ctypes.c_size_t()
            _lib.call("pp_beam_workspace_bytes", ctypes.byref(b))
            ws = self._ws = DeviceArray((int(b.value),), np.uint8)
        return ws

    # -- data ----------------------------------------------------------------------------------
    @property
    def species(self):
        return self._species

    @property
    def x_vec(self):
        if self._host is not None:
            return self._host[0]
        return self._state.positions_aos().to_host()

    @property
    def p_vec(self):
        if self._host is not None:
            return self._host[1]
        w = self._state.velocities_aos().to_host()
        if self._u_state:
            return w / CLIGHT
        b2 = np.minimum(np.sum(w * w, axis=1) / CLIGHT ** 2, 0.9999)
        return w / CLIGHT / np.sqrt(1.0 - b2)[:, None]

    @property
    def v_vec(self):
        return self.to_pusher()[1]

    # -- collective quantities (device reductions) ---------------------------------------------
    def recalculate_all(self):
        """particles.py:685-691: centroid, mean momentum, energy and second moments in one device pass pair."""
        n = self.numpart
        sums = DeviceArray((_lib.MOMENT_SUMS,))
        cen = DeviceArray((_lib.MOMENT_CENTRAL,))
        ws = self._workspace()
        _lib.call("pp_beam_moments", self._args(), int(self._u_state), float(self._species.mass_mev),
                  int(self.z_energy), sums.ptr, cen.ptr, ws.ptr, ws.nbytes, None)
        s, c = sums.to_host(), cen.to_host()
        cnt = s[0] if s[0] > 0 else 1.0
        m = {"n": int(s[0]), "centroid": s[1:4] / cnt, "mean_momentum": s[4:7] / cnt,
             "mean_betagamma": (s[6] if self.z_energy else s[7]) / cnt, "ekin_mean": s[8] / cnt,
             "xp_mean": s[9:12] / cnt, "min": s[12:15], "max": s[15:18],
             "x_std": np.sqrt(c[0:3] / cnt), "xp_std": np.sqrt(c[3:6] / cnt),
             "xxp": c[6] / cnt, "yyp": c[7] / cnt, "xyp": c[8] / cnt, "yxp": c[9] / cnt,
             "ekin_std": np.sqrt(c[10] / cnt)}
        self.numpart = n
        self._m = m
        return m

    def _moments(self):
        return self._m if self._m is not None else self.recalculate_all()

    @property
    def centroid(self):
        return self._moments()["centroid"]

    @property
    def mean_momentum(self):
        return self._moments()["mean_momentum"]

    @property
    def mean_momentum_betagamma(self):
        return self._moments()["mean_betagamma"]

    @property
    def mean_energy_mev(self):
        return self._moments()["ekin_mean"]

    @property
    def rms_energy_spread_mev(self):
        return self._moments()["ekin_std"]

    def get_mean_energy_mev(self):
        return self.mean_energy_mev

    def get_beam_edges(self, mode="1rms"):
        """particles.py:725-759: [xmin, xmax, ymin, ymax, zmin, zmax]."""
        m = self._moments()
        if mode == "1rms":
            lo, hi = m["centroid"] - m["x_std"], m["centroid"] + m["x_std"]
        elif mode == "2rms":
            lo, hi = m["centroid"] - 2.0 * m["x_std"], m["centroid"] + 2.0 * m["x_std"]
        elif mode == "full":
            lo, hi = m["min"], m["max"]
        else:
            raise ValueError(f"Unknown mode: {mode}. Use '1rms', '2rms', or 'full'")
        return np.column_stack((lo, hi)).flatten()

    def get_emittances(self, normalized=True):
        """particles.py:761-788: rms emittances [e_xx', e_yy', e_xy', e_yx'] (times mean beta*gamma if normalized)."""
        m = self._moments()
        sx, sa = m["x_std"], m["xp_std"]
        e = np.sqrt(np.array([(sx[0] * sa[0]) ** 2 - m["xxp"] ** 2, (sx[1] * sa[1]) ** 2 - m["yyp"] ** 2,
                              (sx[0] * sa[1]) ** 2 - m["xyp"] ** 2, (sx[1] * sa[0]) ** 2 - m["yxp"] ** 2]))
        return e * m["mean_betagamma"] if normalized else e

    def get_twiss_parameters(self, report=True):
        """particles.py:790-843: [alpha_x, beta_x, gamma_x, alpha_y, beta_y, gamma_y]; reports the share of
        the beam inside the 4-rms ellipses (counted on the device)."""
        m = self._moments()
        ex, ey, _, _ = self.get_emittances(normalized=False)
        bx, gx = m["x_std"][0] ** 2 / ex, m["xp_std"][0] ** 2 / ex
        ax = np.sqrt(bx * gx - 1.0) if m["xxp"] < 0 else -np.sqrt(bx * gx - 1.0)
        by, gy = m["x_std"][1] ** 2 / ey, m["xp_std"][1] ** 2 / ey
        ay = np.sqrt(by * gy - 1.0) if m["yyp"] < 0 else -np.sqrt(by * gy - 1.0)
        tw = np.array([ax, bx, gx, ay, by, gy])
        cx, cy = self.ellipse_counts(tw, ex, ey)
        self.inside_4rms_percent = (100.0 * cx / max(self.numpart, 1), 100.0 * cy / max(self.numpart, 1))
        if report:
            print(f"4-RMS emittances include {self.inside_4rms_percent[0]:.1f}% and "
                  f"{self.inside_4rms_percent[1]:.1f}% of the beam in x and y direction")
        return tw

    def ellipse_counts(self, twiss, eps_x, eps_y):
        """Particles with gamma x^2 + 2 alpha x x' + beta x'^2 < 4 eps in x and in y."""
        tw = DeviceArray.from_host(np.concatenate([np.asarray(twiss, dtype=np.float64), [eps_x, eps_y]]))
        out = DeviceArray((2,))
        ws = self._workspace()
        _lib.call("pp_beam_ellipse_counts", self._args(), int(self._u_state), tw.ptr, out.ptr, ws.ptr, ws.nbytes,
                  None)
        c = out.to_host()
        return int(c[0]), int(c[1])
