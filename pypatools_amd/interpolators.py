"""Gridded-field interpolators backed by HIP kernels.

Mirrors the factory of src/PyPATools/field_src/interpolators.py:533-610.  Every
backend name the reference accepts ('auto', 'numba', 'fast', 'scipy', 'cupy')
maps to the HIP gather here, plus the explicit name 'hip'.  Semantics follow
NumbaInterpolator (interpolators.py:44-135): linear / bilinear / trilinear on a
rectilinear grid, points outside the grid get ``fill_value``, a point on the
upper face gets the face value.

``DeviceFieldMap`` is the device-resident form of up to six gridded components
sharing one grid (E and B of a field map): one cell lookup serves all of them.
"""
import ctypes

import numpy as np

from . import _lib
from .device import DeviceArray

BACKENDS = ("auto", "hip", "numba", "fast", "scipy", "cupy")


def grid_to_axes(points):
    """Place the 1-, 2- or 3-D grid of an interpolator on the engine's x/y/z axes.

    Field evaluates 1-D fields on pts[:, 2] and 2-D fields on pts[:, :2]
    (field.py:457-489), so a 1-D grid becomes the z axis and a 2-D grid the x
    and y axes; unused axes get a single node.
    """
    pts = [np.ascontiguousarray(p, dtype=np.float64) for p in points]
    one = np.zeros(1)
    if len(pts) == 1:
        return (one, one, pts[0])
    if len(pts) == 2:
        return (pts[0], pts[1], one)
    if len(pts) == 3:
        return tuple(pts)
    raise ValueError(f"Only 1D, 2D, 3D supported. Got {len(pts)}D")


class DeviceFieldMap:
    """A device map of ``ncomp`` components on one grid (pp_map_create)."""

    def __init__(self, points, components, fill_value=0.0, periodic_z=False, fp32=True, coefficients=True):
        axes = grid_to_axes(points)
        shape = tuple(len(a) for a in axes)
        vals = np.empty((len(components),) + shape, dtype=np.float64)
        for c, v in enumerate(components):
            vals[c] = np.asarray(v, dtype=np.float64).reshape(shape)
        flags = 0
        if periodic_z:
            flags |= _lib.MAP_PERIODIC_Z
        if not fp32:
            flags |= _lib.MAP_NO_FP32
        if not coefficients:
            flags |= _lib.MAP_NO_COEF
        h = ctypes.c_void_p()
        vals = np.ascontiguousarray(vals)
        _lib.call("pp_map_create", vals.ctypes.data, len(components), shape[0], shape[1], shape[2],
                  axes[0].ctypes.data, axes[1].ctypes.data, axes[2].ctypes.data, float(fill_value), flags,
                  ctypes.byref(h))
        self.handle = h.value
        self.ncomp = len(components)
        self.shape = shape
        self.periodic_z = periodic_z
        self.fill_value = float(fill_value)

    @property
    def nbytes(self):
        b = ctypes.c_size_t()
        _lib.call("pp_map_bytes", self.handle, ctypes.byref(b))
        return b.value

    def __del__(self):
        if getattr(self, "handle", None):
            try:
                _lib.lib().pp_map_destroy(self.handle)
            except Exception:
                pass
            self.handle = None


def eval_term(term, pts_dev, m, precision, rf_value=1.0, out=None):
    """Evaluate one compiled field term (PPField) at device points (m, 3)."""
    out = out if out is not None else DeviceArray((m, 3))
    _lib.call("pp_field_eval", ctypes.byref(term), _lib.PRECISIONS[precision], pts_dev.ptr, m, float(rf_value),
              out.ptr, None)
    return out


class HipInterpolator:
    """RegularGridInterpolator-compatible callable evaluated by the HIP gather.

    NumbaInterpolator semantics (interpolators.py:273-408): ``method`` must be
    'linear' ('nearest' and higher orders of the 'fast' backend are not on the
    device path), ``bounds_error`` raises for points outside the grid.
    """

    def __init__(self, points, values, method="linear", bounds_error=True, fill_value=np.nan,
                 periodic_z=False):
        if method != "linear":
            raise NotImplementedError(f"Method '{method}' is not implemented by the HIP gather (only 'linear')")
        self.ndim = len(points)
        if self.ndim not in (1, 2, 3):
            raise ValueError(f"Only 1D, 2D, 3D supported. Got {self.ndim}D")
        self.method = method
        self.bounds_error = bounds_error
        self.fill_value = float(fill_value)
        self.grid = tuple(np.ascontiguousarray(p, dtype=np.float64) for p in points)
        self.values = np.ascontiguousarray(values, dtype=np.float64)
        expected = tuple(len(g) for g in self.grid)
        if self.values.shape != expected:
            raise ValueError(f"Values shape {self.values.shape} does not match grid shape {expected}")
        self._bounds = tuple((float(g[0]), float(g[-1])) for g in self.grid)
        self.periodic_z = periodic_z
        self._map = None

    def device_map(self):
        if self._map is None:
            self._map = DeviceFieldMap(self.grid, [self.values, self.values, self.values], self.fill_value,
                                       self.periodic_z)
        return self._map

    def __call__(self, xi, method=None, precision="fp64"):
        xi = np.asarray(xi, dtype=np.float64)
        single = xi.ndim == 1
        if single:
            xi = xi.reshape(1, -1)
        if xi.shape[1] != self.ndim:
            raise ValueError(f"Points have dimension {xi.shape[1]}, expected {self.ndim}")
        if self.bounds_error:
            for a, (lo, hi) in enumerate(self._bounds):
                if np.any(xi[:, a] < lo) or np.any(xi[:, a] > hi):
                    raise ValueError("One or more points are outside the interpolation domain")
        m = xi.shape[0]
        pts = np.zeros((m, 3))
        if self.ndim == 1:
            pts[:, 2] = xi[:, 0]
        else:
            pts[:, :self.ndim] = xi
        term = _lib.PPField()
        term.kind = _lib.FIELD_MAP
        term.map = self.device_map().handle
        term.comp0 = 0
        term.scale1 = 1.0
        if m == 0:
            return np.zeros(0)
        pd = DeviceArray.from_host(pts)
        res = eval_term(term, pd, m, precision).to_host()[:, 0]
        return float(res[0]) if single else res


def get_interpolator(points, values, method="linear", bounds_error=False, fill_value=0.0, backend="auto"):
    """interpolators.py:533-610 with every backend served by the HIP gather."""
    backend = backend.lower()
    if backend not in BACKENDS:
        raise ValueError(f"Unknown backend '{backend}'. Use: 'auto', 'hip', 'numba', 'fast', 'scipy', 'cupy'")
    return HipInterpolator(points, values, method, bounds_error, fill_value)
