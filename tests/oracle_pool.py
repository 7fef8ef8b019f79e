"""The CPU oracle spread over worker processes, for the full-size GPU tests.

Test infrastructure only.  Each job is an oracle ``track_batch`` of a particle
subset through one of the seeded workload maps, returning the last records
(r_hist[-1], v_hist[-1]).  Workers are spawned (fresh interpreters: no GPU
state is inherited) and import numpy, ``oracle`` and the map generators of
``pypatools_amd.workloads`` only, which never load the HIP library.
"""
import multiprocessing as mp
import os
from concurrent.futures import ProcessPoolExecutor

import numpy as np

_FIELDS = {}


def c5_map_scales(k):
    """E and B factors of the k-th of the four resident C5 maps (E * s, B * (2 - s), s = 1 + 0.1 k)."""
    s = 1.0 + 0.1 * k
    return s, 2.0 - s


def _fields(key):
    """Oracle (efield, bfield) for a map key:
    ("c2", n, map_index)  the C2 map on n^3 nodes, periodic in z, with the C5 factors of map_index
                          folded into its node values (map_index None: the plain C2 map);
    ("c3",)               the C3 TM010-like map (static part; the RF factor is applied by ORF)."""
    if key in _FIELDS:
        return _FIELDS[key]
    from oracle.field import OField, OPeriodicZ
    from pypatools_amd.workloads import c2_map, c3_map
    if key[0] == "c2":
        grid, E, B = c2_map(key[1])
        if key[2] is not None:
            se, sb = c5_map_scales(key[2])
            E = {c: se * E[c] for c in "xyz"}
            B = {c: sb * B[c] for c in "xyz"}
        out = (OPeriodicZ(grid, E), OPeriodicZ(grid, B))
    elif key[0] == "c3":
        grid, E, B, _w = c3_map()
        out = (OField.from_arrays(grid, E), OField.from_arrays(grid, B))
    else:
        raise ValueError(key)
    _FIELDS[key] = out
    return out


def track_last(job):
    """One oracle track: job = dict(map, alg, qm, dt, nsteps, r0, v0[, escale, bscale][, rf=(omega, phase)])."""
    from oracle.field import ORF, OScaled
    from oracle.pusher import OraclePusher
    fe, fb = _fields(job["map"])
    if job.get("escale", 1.0) != 1.0:
        fe = OScaled(fe, job["escale"])
    if job.get("bscale", 1.0) != 1.0:
        fb = OScaled(fb, job["bscale"])
    if "rf" in job:
        w, ph = job["rf"]
        fe, fb = ORF(fe, w, ph, "cos"), ORF(fb, w, ph, "sin")
    n = int(job["nsteps"])
    rh, vh, _ = OraclePusher(job["qm"], job["alg"]).track_batch(job["r0"], job["v0"], fe, fb, n, job["dt"],
                                                                max(n, 1))
    return rh[-1], vh[-1]


def workers():
    """Pool size: at most 8, and no more than this host's CPUs."""
    return max(1, min(8, os.cpu_count() or 1))


class OraclePool:
    """A spawned process pool running ``track_last`` jobs; results come back in submission order."""

    def __init__(self, n=None):
        self.ex = ProcessPoolExecutor(max_workers=n or workers(), mp_context=mp.get_context("spawn"))

    def map(self, jobs):
        return list(self.ex.map(track_last, jobs))

    def split(self, job, parts):
        """One job's particles split into `parts` chunks over the pool, reassembled."""
        idx = np.array_split(np.arange(len(job["r0"])), parts)
        subs = [dict(job, r0=job["r0"][i], v0=job["v0"][i]) for i in idx if len(i)]
        res = self.map(subs)
        return np.concatenate([r for r, _ in res]), np.concatenate([v for _, v in res])

    def close(self):
        self.ex.shutdown(wait=True, cancel_futures=True)
