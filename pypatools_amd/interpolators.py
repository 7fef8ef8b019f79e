"""Gridded-field interpolators backed by HIP kernels.

Mirrors the factory of src/PyPATools/field_src/interpolators.py:533-610.  Every
backend name the reference accepts ('auto', 'numba', 'fast', 'scipy', 'cupy')
is served on the device here, plus the explicit name 'hip'.  'linear' follows
NumbaInterpolator (interpolators.py:44-135): linear / bilinear / trilinear on a
rectilinear grid, points outside the grid get ``fill_value``, a point on the
upper face gets the face value.  'nearest', 'cubic' and 'quintic' follow
CoordinateMapper (interpolators.py:415-526) through the device spline.

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

    def __init__(self, points, components, fill_value=0.0, periodic_z=False, fp32=True, coefficients=True,
                 fp64=True):
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
        if not fp64:
            flags |= _lib.MAP_NO_FP64
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


# CoordinateMapper's method -> map_coordinates order (interpolators.py:450-455), and the per-call
# override, which knows no 'quintic' (interpolators.py:482-485: an unknown name keeps the order)
METHOD_ORDERS = {"nearest": 0, "linear": 1, "cubic": 3, "quintic": 5}
CALL_ORDERS = {"nearest": 0, "linear": 1, "cubic": 3}
RGI_METHODS = ("linear", "nearest")        # the RegularGridInterpolator methods served on the device


class DeviceSpline:
    """Device spline interpolator of ``ncomp`` components on one rectilinear grid (pp_spline_create).

    order 0/1/3/5 with map_coordinates semantics after np.interp coordinates; ``nearest_mode``: mode
    'nearest' (else 'constant' with ``cval`` outside the grid); ``rgi``: RegularGridInterpolator's
    'nearest' rule (order 0)."""

    def __init__(self, points, components, order, nearest_mode=False, rgi=False, cval=0.0):
        axes = [np.ascontiguousarray(p, dtype=np.float64) for p in points]
        ndim = len(axes)
        if ndim not in (1, 2, 3):
            raise ValueError(f"Only 1D, 2D, 3D supported. Got {ndim}D")
        shape = tuple(len(a) for a in axes)
        vals = np.empty((len(components),) + shape, dtype=np.float64)
        for c, v in enumerate(components):
            vals[c] = np.asarray(v, dtype=np.float64).reshape(shape)
        flags = (_lib.SPLINE_NEAREST_MODE if nearest_mode else 0) | (_lib.SPLINE_RGI if rgi else 0)
        sh = np.array(shape, dtype=np.int64)
        ax = axes + [None] * (3 - ndim)
        h = ctypes.c_void_p()
        _lib.call("pp_spline_create", vals.ctypes.data, len(components), ndim, sh.ctypes.data,
                  *[a.ctypes.data if a is not None else None for a in ax], int(order), flags, float(cval),
                  ctypes.byref(h))
        self.handle = h.value
        self.ncomp = len(components)
        self.ndim = ndim
        self.order = order

    def eval_device(self, pts_dev, m, out, stride=None, col0=0, stream=None):
        """Values at device points pts_dev (row i: coordinates at [i*stride + col0 ...]) into out (m, ncomp)."""
        _lib.call("pp_spline_eval", self.handle, pts_dev.ptr, int(m), int(stride or self.ndim), int(col0), out.ptr,
                  stream)
        return out

    def __del__(self):
        if getattr(self, "handle", None):
            try:
                _lib.lib().pp_spline_destroy(self.handle)
            except Exception:
                pass
            self.handle = None


class HipInterpolator:
    """Interpolator callable evaluated on the device, with the semantics of the reference backend
    it stands in for (interpolators.py:533-610):

    * 'auto', 'hip', 'fast', 'numba', 'cupy' (numba and cupy are absent, so the reference falls back
      to 'fast'): CoordinateMapper (interpolators.py:415-526).  'linear' runs the trilinear gather
      (NumbaInterpolator semantics, equal to map_coordinates order 1 to rounding); 'nearest',
      'cubic', 'quintic' run the device spline (np.interp coordinates, map_coordinates orders 0, 3,
      5, mode 'nearest' when ``bounds_error`` else 'constant' with ``fill_value``).  A call may
      override the method with 'nearest', 'linear' or 'cubic'.
    * 'scipy': RegularGridInterpolator for 'linear' and 'nearest' (per-call method too); its
      spline methods raise NotImplementedError.

    ``bounds_error`` raises ValueError for points outside the grid.
    """

    def __init__(self, points, values, method="linear", bounds_error=True, fill_value=np.nan,
                 periodic_z=False, backend="hip"):
        self.rgi = backend == "scipy"
        if self.rgi:
            if method not in RGI_METHODS:
                raise NotImplementedError(f"Method '{method}' of the 'scipy' backend is not implemented on the "
                                          f"device (only {RGI_METHODS})")
            if fill_value is None:
                raise NotImplementedError("fill_value=None (extrapolation) is not implemented on the device")
            self.order = 0 if method == "nearest" else 1
        else:
            if method not in METHOD_ORDERS:
                raise ValueError(f"Method '{method}' not supported")
            self.order = METHOD_ORDERS[method]
        self.ndim = len(points)
        if self.ndim not in (1, 2, 3):
            raise ValueError(f"Only 1D, 2D, 3D supported. Got {self.ndim}D")
        self.method = method
        self.backend = backend
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
        self._splines = {}

    def device_map(self):
        if self._map is None:
            self._map = DeviceFieldMap(self.grid, [self.values, self.values, self.values], self.fill_value,
                                       self.periodic_z)
        return self._map

    def device_spline(self, order):
        """The device spline of `order` for these values (cached)."""
        if self.periodic_z:
            raise NotImplementedError("periodic_z maps are interpolated linearly only")
        if order not in self._splines:
            self._splines[order] = DeviceSpline(self.grid, [self.values], order, nearest_mode=self.bounds_error,
                                                rgi=self.rgi and order == 0,
                                                cval=0.0 if self.bounds_error else self.fill_value)
        return self._splines[order]

    def call_order(self, method=None):
        """The order a call evaluates: the per-call method of interpolators.py:482-485 (or RGI's)."""
        if method is None:
            return self.order
        if self.rgi:
            if method not in RGI_METHODS:
                raise NotImplementedError(f"Method '{method}' of the 'scipy' backend is not implemented on the device")
            return 0 if method == "nearest" else 1
        return CALL_ORDERS.get(method, self.order)

    def __call__(self, xi, method=None, precision="fp64"):
        xi = np.asarray(xi, dtype=np.float64)
        single = xi.ndim == 1
        if single:
            xi = xi.reshape(1, -1)
        if xi.shape[1] != self.ndim:
            raise ValueError(f"Points have dimension {xi.shape[1]}, expected {self.ndim}")
        order = self.call_order(method)
        if self.bounds_error:
            for a, (lo, hi) in enumerate(self._bounds):
                if np.any(xi[:, a] < lo) or np.any(xi[:, a] > hi):
                    raise ValueError("One or more points are outside the interpolation domain")
        m = xi.shape[0]
        if m == 0:
            return np.zeros(0)
        if order != 1:
            pd = DeviceArray.from_host(np.ascontiguousarray(xi))
            out = DeviceArray((m, 1))
            res = self.device_spline(order).eval_device(pd, m, out).to_host()[:, 0]
            return float(res[0]) if single else res
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
        pd = DeviceArray.from_host(pts)
        res = eval_term(term, pd, m, precision).to_host()[:, 0]
        return float(res[0]) if single else res


def get_interpolator(points, values, method="linear", bounds_error=False, fill_value=0.0, backend="auto"):
    """interpolators.py:533-610 with every backend served on the device (see HipInterpolator)."""
    backend = backend.lower()
    if backend not in BACKENDS:
        raise ValueError(f"Unknown backend '{backend}'. Use: 'auto', 'hip', 'numba', 'fast', 'scipy', 'cupy'")
    return HipInterpolator(points, values, method, bounds_error, fill_value, backend=backend)
