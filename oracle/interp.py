"""Gridded field gather (linear / bilinear / trilinear).  Test infrastructure only.

``interp_nd`` restates the NumbaInterpolator kernels of
src/PyPATools/field_src/interpolators.py vectorised over points:

* out of ``[g[0], g[-1]]`` on any axis -> ``fill_value``      (:94-97, :196-197, :213-214)
* cell index = binary search clamped to ``[0, n-2]``           (:45-60, :100-106)
  (== ``searchsorted(g, x, 'right') - 1`` clipped; a point on the upper face
  lands in the last cell with weight 1)
* weight ``t = (x - g[i]) / (g[i+1] - g[i])`` clamped to [0, 1] (:109-115)
* nested lerps ``a*(1-t) + b*t``: x first, then y, then z      (:117-135)

``interp_coordinate_mapper`` restates CoordinateMapper (interpolators.py:415-526):
np.interp to fractional indices (left=-1, right=n) then
scipy.ndimage.map_coordinates(order, mode='constant', cval=fill) -- order 0, 1,
3 or 5 for the methods 'nearest', 'linear', 'cubic', 'quintic'.  This is the
path the reference takes when numba is absent (interpolators.py:567-571).  Its
arithmetic is the library routine itself (scipy 1.15 here); tests pin it further
by the interpolation property (node values are reproduced), the rounding rule of
order 0 and the exactness of order 1 on linear data.
``interp_rgi_nearest`` is the 'scipy' backend's RegularGridInterpolator
'nearest' (the library routine).
"""
import numpy as np
from scipy.ndimage import map_coordinates


def cell_and_weight(g, x):
    """Per-axis cell index and clamped weight, interpolators.py:45-60, 100-115."""
    g = np.asarray(g, dtype=np.float64)
    n = len(g)
    i = np.searchsorted(g, x, side="right") - 1
    i = np.clip(i, 0, n - 2)
    t = (x - g[i]) / (g[i + 1] - g[i])
    t = np.maximum(0.0, np.minimum(t, 1.0))
    return i, t


def inside(g, x):
    return ~((x < g[0]) | (x > g[-1]))


def _lerp(a, b, t):
    return a * (1.0 - t) + b * t


def interp_nd(points, values, xi, fill_value=0.0):
    """NumbaInterpolator.__call__ batch semantics for 1-D/2-D/3-D grids.

    points: tuple of 1-D node arrays; values: ndarray of matching shape;
    xi: (M, ndim) query points.  Returns (M,) float64.
    """
    xi = np.atleast_2d(np.asarray(xi, dtype=np.float64))
    values = np.ascontiguousarray(values, dtype=np.float64)
    nd = len(points)
    ok = np.ones(xi.shape[0], dtype=bool)
    idx, wts = [], []
    for a in range(nd):
        g = np.asarray(points[a], dtype=np.float64)
        ok &= inside(g, xi[:, a])
        i, t = cell_and_weight(g, xi[:, a])
        idx.append(i)
        wts.append(t)
    if nd == 1:
        (i,), (t,) = idx, wts
        res = values[i] * (1.0 - t) + values[i + 1] * t  # interpolators.py:205
    elif nd == 2:
        (i, j), (tx, ty) = idx, wts
        c0 = _lerp(values[i, j], values[i + 1, j], tx)      # :233
        c1 = _lerp(values[i, j + 1], values[i + 1, j + 1], tx)  # :234
        res = _lerp(c0, c1, ty)                             # :236
    elif nd == 3:
        (i, j, k), (tx, ty, tz) = idx, wts
        c00 = _lerp(values[i, j, k], values[i + 1, j, k], tx)              # :127
        c01 = _lerp(values[i, j, k + 1], values[i + 1, j, k + 1], tx)      # :128
        c10 = _lerp(values[i, j + 1, k], values[i + 1, j + 1, k], tx)      # :129
        c11 = _lerp(values[i, j + 1, k + 1], values[i + 1, j + 1, k + 1], tx)  # :130
        c0 = _lerp(c00, c10, ty)                                           # :132
        c1 = _lerp(c01, c11, ty)                                           # :133
        res = _lerp(c0, c1, tz)                                            # :135
    else:
        raise ValueError(f"Only 1D, 2D, 3D supported. Got {nd}D")
    return np.where(ok, res, float(fill_value))


def interp_coordinate_mapper(points, values, xi, fill_value=0.0, order=1, bounds_error=False):
    """CoordinateMapper.__call__ (interpolators.py:469-520): np.interp to fractional node indices
    (left=-1, right=n), then map_coordinates of `order` (method 'nearest' 0, 'linear' 1, 'cubic' 3,
    'quintic' 5, :450-455) in mode 'constant' with cval=fill_value, or mode 'nearest' with cval 0
    when bounds_error (:504-510; the bounds check itself, :523-526, is the caller's)."""
    xi = np.atleast_2d(np.asarray(xi, dtype=np.float64))
    coords = np.empty((xi.shape[0], len(points)))
    for a, g in enumerate(points):
        g = np.asarray(g)
        coords[:, a] = np.interp(xi[:, a], g, np.arange(len(g)), left=-1, right=len(g))
    mode = "nearest" if bounds_error else "constant"
    cval = 0.0 if bounds_error else fill_value
    return map_coordinates(np.asarray(values, dtype=np.float64), coords.T, order=order, mode=mode, cval=cval)


def interp_rgi_nearest(points, values, xi, fill_value=0.0):
    """The 'scipy' backend's method 'nearest' (interpolators.py:585): RegularGridInterpolator."""
    from scipy.interpolate import RegularGridInterpolator
    rgi = RegularGridInterpolator(tuple(np.asarray(p, dtype=np.float64) for p in points),
                                  np.asarray(values, dtype=np.float64), method="nearest", bounds_error=False,
                                  fill_value=fill_value)
    return rgi(np.atleast_2d(np.asarray(xi, dtype=np.float64)))
