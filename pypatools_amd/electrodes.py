"""Electrode surfaces as triangle meshes, tested on the device.

``TriangleElectrodes`` speaks the reference's electrode protocol
(``segment_intersects_surface(r_old, r_new) -> {'hit_mask': ...}``, used by
Pusher.track_batch, src/PyPATools/pusher.py:982-990), so it can be passed as a
Pusher's ``electrode_assembly``.  With it the engine keeps the loss test on the
device: after each step the kernel tests every live particle's segment against
the mesh's bounding-volume hierarchy and clears its alive flag on a hit, with no
host round trip per step.

A segment hits when it meets a triangle on the closed intervals (edges and end
points count); a segment lying in a triangle's plane does not.

The same meshes bound the space-charge solve (space_charge.SpaceChargeSolver):
closed surfaces are grounded solid electrodes.  `box_mesh`, `cylinder_mesh` and
`sphere_mesh` build (vertices, triangles) arrays without touching the device.
"""
import ctypes

import numpy as np

from . import _lib
from .device import DeviceArray


def box_mesh(lo, hi):
    """Closed axis-aligned box surface, 12 triangles (each face split along its lo-lo / hi-hi diagonal)."""
    lo, hi = np.asarray(lo, dtype=np.float64), np.asarray(hi, dtype=np.float64)
    v = np.array([[lo[0] if i & 1 == 0 else hi[0], lo[1] if i & 2 == 0 else hi[1], lo[2] if i & 4 == 0 else hi[2]]
                  for i in range(8)])
    quads = [(0, 1, 3, 2), (4, 6, 7, 5), (0, 4, 5, 1), (2, 3, 7, 6), (0, 2, 6, 4), (1, 5, 7, 3)]
    tris = [t for a, b, c, d in quads for t in ((a, b, c), (a, c, d))]
    return v, np.array(tris, dtype=np.int32)


def cylinder_mesh(radius, z0, z1, n_phi=64, n_z=16):
    """Open tube of the given radius along z (a beam pipe), 2 n_phi n_z triangles."""
    phi = np.linspace(0.0, 2 * np.pi, n_phi, endpoint=False)
    zs = np.linspace(z0, z1, n_z + 1)
    v = np.array([[radius * np.cos(p), radius * np.sin(p), z] for z in zs for p in phi])
    tris = []
    for k in range(n_z):
        for j in range(n_phi):
            a, b = k * n_phi + j, k * n_phi + (j + 1) % n_phi
            c, d = a + n_phi, b + n_phi
            tris += [(a, b, d), (a, d, c)]
    return v, np.array(tris, dtype=np.int32)


def sphere_mesh(center, radius, subdivisions=2):
    """Closed icosphere (20 * 4**subdivisions triangles), outward-facing, vertices on the sphere."""
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    v = [np.array(p, dtype=np.float64) / np.linalg.norm(p) for p in v]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6),
         (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10),
         (8, 6, 7), (9, 8, 1)]
    for _ in range(subdivisions):
        mid = {}

        def m(a, b):
            k = (min(a, b), max(a, b))
            if k not in mid:
                p = v[a] + v[b]
                v.append(p / np.linalg.norm(p))
                mid[k] = len(v) - 1
            return mid[k]

        f = [g for a, b, c in f for g in ((a, m(a, b), m(c, a)), (b, m(b, c), m(a, b)), (c, m(c, a), m(b, c)),
                                          (m(a, b), m(b, c), m(c, a)))]
    return np.asarray(center, dtype=np.float64) + radius * np.array(v), np.array(f, dtype=np.int32)


class TriangleElectrodes:
    def __init__(self, vertices, triangles):
        self.vertices = np.ascontiguousarray(vertices, dtype=np.float64).reshape(-1, 3)
        self.triangles = np.ascontiguousarray(triangles, dtype=np.int32).reshape(-1, 3)
        h = ctypes.c_void_p()
        _lib.call("pp_mesh_create", self.vertices.ctypes.data, len(self.vertices), self.triangles.ctypes.data,
                  len(self.triangles), ctypes.byref(h))
        self.handle = h.value

    def __del__(self):
        h = getattr(self, "handle", None)
        if h:
            try:
                _lib.lib().pp_mesh_destroy(ctypes.c_void_p(h))
            except Exception:
                pass
            self.handle = None

    @property
    def n_triangles(self):
        return len(self.triangles)

    def info(self):
        nt, nn = ctypes.c_int64(), ctypes.c_int64()
        _lib.call("pp_mesh_info", self.handle, ctypes.byref(nt), ctypes.byref(nn))
        return {"triangles": nt.value, "nodes": nn.value}

    # -- the reference's protocol (host arrays in and out) ------------------------------------
    def segment_intersects_surface(self, r_old, r_new):
        r0 = np.ascontiguousarray(r_old, dtype=np.float64).reshape(-1, 3)
        r1 = np.ascontiguousarray(r_new, dtype=np.float64).reshape(-1, 3)
        n = len(r0)
        if n == 0:
            return {"hit_mask": np.zeros(0, dtype=bool)}
        a, b = DeviceArray.from_host(np.ascontiguousarray(r0.T)), DeviceArray.from_host(np.ascontiguousarray(r1.T))
        hit = DeviceArray((n,), np.uint8)
        self.collide_device((a.view_ptr(0), a.view_ptr(n), a.view_ptr(2 * n)),
                            (b.view_ptr(0), b.view_ptr(n), b.view_ptr(2 * n)), n, None, hit.ptr)
        return {"hit_mask": hit.to_host().astype(bool)}

    # -- device form, used by the tracking engine ----------------------------------------------
    def collide_device(self, old_xyz, new_xyz, n, alive_ptr, hit_ptr=None, stream=None):
        _lib.call("pp_collide_segments", self.handle, *old_xyz, *new_xyz, n, alive_ptr, hit_ptr, stream)

    # -- simple meshes -----------------------------------------------------------------------
    @classmethod
    def box(cls, lo, hi):
        """Closed axis-aligned box surface, 12 triangles."""
        return cls(*box_mesh(lo, hi))

    @classmethod
    def cylinder(cls, radius, z0, z1, n_phi=64, n_z=16):
        """Open tube of the given radius along z (a beam pipe), 2 n_phi n_z triangles."""
        return cls(*cylinder_mesh(radius, z0, z1, n_phi, n_z))

    @classmethod
    def sphere(cls, center, radius, subdivisions=2):
        """Closed icosphere, 20 * 4**subdivisions triangles."""
        return cls(*sphere_mesh(center, radius, subdivisions))

    @classmethod
    def merge(cls, *meshes):
        vs, ts, off = [], [], 0
        for m in meshes:
            vs.append(m.vertices)
            ts.append(m.triangles + off)
            off += len(m.vertices)
        return cls(np.concatenate(vs), np.concatenate(ts))
