"""Pusher: the reference's particle-pusher API on the HIP engine.

Same surface as src/PyPATools/pusher.py:430-1037: algorithm chosen by name
(``ALGORITHMS``, the reference's seven names; unknown names raise the same
ValueError with the same text), ``push`` /
``push_batch`` / ``track`` / ``track_batch`` / ``change_algorithm``, host
float64 (M, 3) arrays in and out.  All arithmetic runs in the HIP kernels of
libpypat_hip.so; ``use_numba`` is accepted and ignored.

Additions: algorithms 'boris_rel' (relativistic Boris) and 'vay' (Vay 2008),
both staggered like 'boris', listed in ``EXTENSION_ALGORITHMS`` and accepted
by name beside ``ALGORITHMS``; ``precision`` ('fp64' = bitwise the reference's
FP64 NumPy arithmetic, 'fp64_fast', 'fp32_mixed'); ``t0`` for time-dependent
fields (RFField); per-particle ``dt`` arrays for the RK4 family
(pusher.py:815-816).
"""
import warnings

import numpy as np

from . import _lib
from .engine import EXTENSION_ALGORITHMS, REFERENCE_ALGORITHMS, Tracker


class Pusher:
    """Particle pusher for tracking charged particles through EM fields."""

    ALGORITHMS = list(REFERENCE_ALGORITHMS)           # pusher.py:465-466
    EXTENSION_ALGORITHMS = list(EXTENSION_ALGORITHMS)

    def __init__(self, ion, algorithm="boris", use_numba=True, electrode_assembly=None, precision="fp64",
                 block_size=0, sort_every=None):
        self.ion = ion
        self.q_over_m = ion.q_over_m
        algorithm = algorithm.lower()
        self._check_name(algorithm)
        # pusher.py:480-489: a leapfrog request for an ion that carries an energy above 100 keV
        if algorithm == "leapfrog" and hasattr(ion, "energy_mev"):
            if ion.energy_mev > 0.1:
                warnings.warn("Leapfrog is non-relativistic. Using 'vay_rel' instead for better relativistic "
                              "accuracy.", UserWarning)
                algorithm = "vay_rel"
        if precision not in _lib.PRECISIONS:
            raise ValueError(f"Unknown precision '{precision}'. Must be one of {list(_lib.PRECISIONS)}")
        self.algorithm = algorithm
        self.use_numba = use_numba
        self.relativistic = algorithm.endswith("_rel") or algorithm == "vay"
        self.elec_assy = electrode_assembly
        self.precision = precision
        self.block_size = block_size
        self.sort_every = sort_every

    def _check_name(self, algorithm):
        """pusher.py:475-478: the message lists ALGORITHMS (the extensions are accepted, not listed)."""
        if algorithm not in self.ALGORITHMS and algorithm not in self.EXTENSION_ALGORITHMS:
            raise ValueError(f"Unknown algorithm '{algorithm}'. Must be one of {self.ALGORITHMS}")

    def _tracker(self):
        kw = {} if self.sort_every is None else {"sort_every": self.sort_every}
        return Tracker(self.algorithm, self.precision, self.q_over_m, self.block_size, **kw)

    # ---- single particle ------------------------------------------------------------
    def push(self, r, v, efield, bfield, dt, t=0.0):
        """One step of one particle (pusher.py:508-565); r, v: (3,)."""
        r = np.asarray(r, dtype=np.float64)
        v = np.asarray(v, dtype=np.float64)
        rn, vn = self._one_step(r.reshape(1, 3), v.reshape(1, 3), efield, bfield, dt, t)
        return rn[0], vn[0]

    def track(self, r0, v0, efield, bfield, nsteps, dt, rec_every_n_steps=1, t0=0.0):
        """Track one particle (pusher.py:620-693); zero-initialised (n_records, 3) histories."""
        rh, vh, _ = self._tracker().track(np.asarray(r0, dtype=np.float64).reshape(1, 3),
                                          np.asarray(v0, dtype=np.float64).reshape(1, 3), efield, bfield,
                                          nsteps, dt, rec_every_n_steps, batch=False, t0=t0)
        return rh[:, 0, :], vh[:, 0, :]

    # ---- batches ------------------------------------------------------------------------
    def push_batch(self, r_array, v_array, efield, bfield, dt, t=0.0):
        """One step of M particles (pusher.py:699-801); returns new arrays."""
        return self._one_step(np.asarray(r_array, dtype=np.float64), np.asarray(v_array, dtype=np.float64),
                              efield, bfield, dt, t)

    def _one_step(self, r, v, efield, bfield, dt, t):
        dt_pp = None
        if np.ndim(dt) == 1:
            if self.algorithm not in ("rk4", "rk4_rel"):
                raise ValueError("per-particle dt is supported by 'rk4' and 'rk4_rel' (pusher.py:815-816)")
            dt_pp, dt = np.asarray(dt, dtype=np.float64), 0.0
        return self._tracker().single_step(r, v, efield, bfield, dt, t, dt_pp)

    def track_batch(self, r0_array, v0_array, efield, bfield, nsteps, dt, rec_every_n_steps=1, verbose=False,
                    t0=0.0):
        """Track M particles (pusher.py:894-1012); returns (r_hist, v_hist, active)."""
        return self._tracker().track(r0_array, v0_array, efield, bfield, nsteps, dt, rec_every_n_steps,
                                     batch=True, t0=t0, electrode_assembly=self.elec_assy, verbose=verbose)

    # ---- utilities ----------------------------------------------------------------------
    def change_algorithm(self, algorithm):
        """pusher.py:1018-1033 (like the reference, the electrode assembly is not carried over)."""
        algorithm = algorithm.lower()
        self._check_name(algorithm)
        self.__init__(self.ion, algorithm=algorithm, use_numba=self.use_numba, precision=self.precision,
                      block_size=self.block_size, sort_every=self.sort_every)

    def __repr__(self):
        """pusher.py:1035-1037; a precision other than the default 'fp64' is appended."""
        extra = "" if self.precision == "fp64" else f", precision='{self.precision}'"
        return (f"Pusher(ion={self.ion.name}, algorithm='{self.algorithm}', "
                f"relativistic={self.relativistic}, numba={self.use_numba}{extra})")
