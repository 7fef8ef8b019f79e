"""Brute-force segment / triangle-mesh intersection (test infrastructure only).

The device engine tests particle steps r_old -> r_new against electrode meshes,
the role of electrode_assembly.segment_intersects_surface in Pusher.track_batch
(src/PyPATools/pusher.py:982-990; the mesh package behind it, py_electrodes, is
not part of the reference tree).  This oracle tests every segment against every
triangle with the Moller-Trumbore form on closed intervals, operation for
operation as the device kernel does it.
"""
import numpy as np


def segment_hits(vertices, triangles, r0, r1):
    """hit[i] = segment r0[i] -> r1[i] meets any triangle (edges and end points included)."""
    v = np.asarray(vertices, dtype=np.float64)
    t = np.asarray(triangles, dtype=np.int64)
    r0 = np.asarray(r0, dtype=np.float64).reshape(-1, 3)
    r1 = np.asarray(r1, dtype=np.float64).reshape(-1, 3)
    v0 = v[t[:, 0]]
    e1 = v[t[:, 1]] - v0
    e2 = v[t[:, 2]] - v0
    hit = np.zeros(len(r0), dtype=bool)
    for i in range(len(r0)):
        p = r0[i]
        d = r1[i] - p
        px = d[1] * e2[:, 2] - d[2] * e2[:, 1]
        py = d[2] * e2[:, 0] - d[0] * e2[:, 2]
        pz = d[0] * e2[:, 1] - d[1] * e2[:, 0]
        det = e1[:, 0] * px + e1[:, 1] * py + e1[:, 2] * pz
        with np.errstate(divide="ignore", invalid="ignore"):
            inv = 1.0 / det
            sx, sy, sz = p[0] - v0[:, 0], p[1] - v0[:, 1], p[2] - v0[:, 2]
            u = (sx * px + sy * py + sz * pz) * inv
            qx = sy * e1[:, 2] - sz * e1[:, 1]
            qy = sz * e1[:, 0] - sx * e1[:, 2]
            qz = sx * e1[:, 1] - sy * e1[:, 0]
            w = (d[0] * qx + d[1] * qy + d[2] * qz) * inv
            s = (e2[:, 0] * qx + e2[:, 1] * qy + e2[:, 2] * qz) * inv
        ok = (det != 0.0) & (u >= 0.0) & (u <= 1.0) & (w >= 0.0) & (u + w <= 1.0) & (s >= 0.0) & (s <= 1.0)
        hit[i] = bool(np.any(ok))
    return hit


class OracleElectrodes:
    """The reference's electrode protocol over the brute-force test."""

    def __init__(self, vertices, triangles):
        self.vertices, self.triangles = vertices, triangles

    def segment_intersects_surface(self, r_old, r_new):
        return {"hit_mask": segment_hits(self.vertices, self.triangles, r_old, r_new)}
