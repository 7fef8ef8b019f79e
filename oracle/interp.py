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
scipy.ndimage.map_coordinates(order=1, mode='constant', cval=fill).  This is
the path the reference takes when numba is absent (interpolators.py:567-571).
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


def interp_coordinate_mapper(points, values, xi, fill_value=0.0):
    """CoordinateMapper.__call__ with bounds_error=False, order=1 (interpolators.py:469-520)."""
    xi = np.atleast_2d(np.asarray(xi, dtype=np.float64))
    coords = np.empty((xi.shape[0], len(points)))
    for a, g in enumerate(points):
        g = np.asarray(g)
        coords[:, a] = np.interp(xi[:, a], g, np.arange(len(g)), left=-1, right=len(g))
    return map_coordinates(np.asarray(values), coords.T, order=1, mode="constant", cval=fill_value)
