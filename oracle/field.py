"""Field objects of the reference, restated.  Test infrastructure only.

* ``OField``      -- Field.__call__ dispatch by dim (field.py:213-244):
                     0-D ``scaling * value * ones(M)`` (:441-455),
                     1-D on pts[:, 2] (:457-472), 2-D on pts[:, :2] (:474-489),
                     3-D on pts (:491-504); each component ``scaling * interp``.
                     ``from_arrays`` drops axes of length 1 (:345-387).
* ``OScaled``     -- ScaledField ``scale * F + offset`` (field.py:661-690).
* ``OComposite``  -- CompositeField NumPy path: zeros, then ``+= w * F`` (:626-655).
* ``OZero``       -- Field.zero(): exactly 0 everywhere inside +-1e20 (:389-407).

Engine extensions (no reference symbol; pinned by closed forms):
* ``OPeriodicZ``  -- a wrapping callable: z -> z0 + mod(z - z0, L) before the
                     gather, over a map whose z axis is extended by one node
                     (z0 + L) that repeats node 0.
* ``ORF``         -- time-dependent ``F(x) * cos(omega t + phase)`` (or sin).
"""
import numpy as np

from .interp import interp_coordinate_mapper, interp_nd


class OField:
    def __init__(self, dim=0, const=None, points=None, values=None, scaling=1.0, interp="numba", order=1):
        self.dim = dim
        self.const = const
        self.points = points
        self.values = values
        self.scaling = scaling
        # 'numba': NumbaInterpolator semantics (canonical); 'coordmapper': the
        # CoordinateMapper path the reference takes without numba, of spline `order`
        # (Field(method='nearest' | 'linear' | 'cubic' | 'quintic'), interpolators.py:450-455).
        if interp == "numba":
            self._interp = interp_nd
        else:
            self._interp = lambda p, v, q, fill: interp_coordinate_mapper(p, v, q, fill, order=order)

    @classmethod
    def constant(cls, fx, fy, fz, scaling=1.0):
        return cls(dim=0, const={"x": fx, "y": fy, "z": fz}, scaling=scaling)

    @classmethod
    def from_arrays(cls, grid, values, scaling=1.0, interp="numba", order=1):
        """field.py:369-387: dim = number of axes with more than one node."""
        pts = tuple(np.asarray(grid[k], dtype=np.float64) for k in "xyz" if k in grid and len(grid[k]) > 1)
        vals = {c: np.asarray(values[c], dtype=np.float64).reshape(tuple(len(p) for p in pts)) for c in values}
        return cls(dim=len(pts), points=pts, values=vals, scaling=scaling, interp=interp, order=order)

    def __call__(self, pts):
        pts = np.atleast_2d(np.asarray(pts, dtype=np.float64))
        if pts.shape[1] != 3:
            raise ValueError(f"Points must have shape (M, 3), got {pts.shape}")
        M = len(pts)
        if self.dim == 0:
            cols = [self.scaling * self.const[c] * np.ones(M) for c in "xyz"]
        else:
            if self.dim == 1:
                q = pts[:, 2:3]
            elif self.dim == 2:
                q = pts[:, :2]
            else:
                q = pts
            cols = [self.scaling * self._interp(self.points, self.values[c], q, 0.0) for c in "xyz"]
        return np.column_stack(cols)


class OZero:
    def __call__(self, pts):
        pts = np.atleast_2d(pts)
        return np.zeros((len(pts), 3))


class OScaled:
    def __init__(self, field, scale=1.0, offset=0.0):
        self.field, self.scale, self.offset = field, scale, offset

    def __call__(self, pts):
        return self.scale * self.field(pts) + self.offset


class OComposite:
    def __init__(self, fields, weights):
        self.fields = fields
        self.weights = np.array(weights, dtype=np.float64)

    def __call__(self, pts):
        pts = np.atleast_2d(pts)
        total = np.zeros((len(pts), 3), dtype=np.float64)
        for f, w in zip(self.fields, self.weights):
            total += w * f(pts)
        return total


def periodic_extend(gz, values_zlast_axis):
    """Extend a uniform z axis by one node z0 + L (L = nz * dz) repeating node 0.

    Returns (gz_ext, L, values_ext).  ``values_zlast_axis`` has z as its last axis.
    """
    gz = np.asarray(gz, dtype=np.float64)
    nz = len(gz)
    dz = (gz[-1] - gz[0]) / (nz - 1)
    L = nz * dz
    gz_ext = np.append(gz, gz[0] + L)
    v = np.asarray(values_zlast_axis, dtype=np.float64)
    v_ext = np.concatenate([v, v[..., :1]], axis=-1)
    return gz_ext, L, v_ext


def wrap_z(pts, z0, L):
    """z -> z0 + mod(z - z0, L) (numpy remainder: result has the sign of L)."""
    out = np.array(pts, dtype=np.float64, copy=True)
    out[:, 2] = z0 + np.mod(out[:, 2] - z0, L)
    return out


class OPeriodicZ:
    """3-D map periodic along z: grid/values as for OField.from_arrays (3-D)."""

    def __init__(self, grid, values, scaling=1.0):
        gz_ext, L, _ = periodic_extend(grid["z"], np.zeros(len(grid["z"])))
        self.z0 = float(np.asarray(grid["z"], dtype=np.float64)[0])
        self.L = L
        g2 = {"x": grid["x"], "y": grid["y"], "z": gz_ext}
        v2 = {c: periodic_extend(grid["z"], values[c])[2] for c in values}
        self.inner = OField.from_arrays(g2, v2, scaling=scaling)

    def __call__(self, pts):
        return self.inner(wrap_z(np.atleast_2d(pts), self.z0, self.L))


class ORF:
    """Time-dependent field F(x) * cos(omega t + phase)  (kind='cos') or sin."""

    time_dependent = True

    def __init__(self, field, omega, phase, kind="cos"):
        self.field, self.omega, self.phase, self.kind = field, omega, phase, kind

    def factor(self, t):
        arg = self.omega * t + self.phase
        return np.cos(arg) if self.kind == "cos" else np.sin(arg)

    def __call__(self, pts, t):
        return self.field(pts) * self.factor(t)
