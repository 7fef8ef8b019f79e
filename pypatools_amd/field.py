"""Electromagnetic field objects with device-side evaluation.

Same surface as src/PyPATools/field.py: ``FieldBase`` (:68-98), ``Field``
(:101-559) with ``__call__`` (:213-244), ``from_arrays`` (:345-387), ``zero``
(:389-407), ``scaling``/``label``/``dim``/``metadata`` and the field algebra
(:260-278); ``CompositeField`` (:595-658), ``ScaledField`` (:661-690),
``FieldMap`` (:697-755) and the coordinate helpers (:762-824).

Evaluation runs on the GPU: a Field compiles to a device *term* (constant or
gridded map + its scaling chain) that the push kernels read directly, and
``__call__`` runs the same gather kernel on the given points.  Expressions the
kernels cannot read directly (sums of fields on different grids, nested
algebra) are evaluated term by term on the device and combined there.

Extensions (no reference symbol): ``periodic_z`` maps, ``RFField`` time
dependence ``F(x) cos(omega t + phase)``, and the ``precision`` of evaluation.
"""
import os
import warnings
from abc import ABC, abstractmethod

import numpy as np

from . import _lib
from .device import DeviceArray
from .interpolators import BACKENDS, DeviceFieldMap, DeviceSpline, HipInterpolator, eval_term, get_interpolator

UNIT_SCALES = {
    "m": 1.0, "cm": 0.01, "mm": 0.001,
    "Tesla": 1.0, "T": 1.0, "Gauss": 1e-4, "kGauss": 0.1,
    "V/m": 1.0, "V/cm": 100.0, "V/mm": 1000.0,
}


class FieldBase(ABC):
    """Abstract base class for all field types (field.py:68-98)."""

    @abstractmethod
    def __call__(self, pts):
        """Evaluate the field at points (M, 3) -> (M, 3)."""

    def gradient(self, pts):
        raise NotImplementedError("Gradient not implemented for this field type")

    # field algebra, field.py:260-278 (available on every field here)
    def __add__(self, other):
        if isinstance(other, (int, float)):
            return ScaledField(self, offset=other)
        return CompositeField([self, other], [1.0, 1.0])

    def __mul__(self, scalar):
        return ScaledField(self, scale=scalar)

    def __rmul__(self, scalar):
        return self.__mul__(scalar)

    def __sub__(self, other):
        if isinstance(other, (int, float)):
            return ScaledField(self, offset=-other)
        return CompositeField([self, other], [1.0, -1.0])

    # device evaluation of an arbitrary field expression at device points
    def _eval_device(self, pts_dev, m, precision, t=None):
        return evaluate_on_device(self, pts_dev, m, precision, t)


_FIELD_KWARGS = ("label", "dim", "field", "scaling", "units", "debug", "method", "interpolator_backend",
                 "periodic_z", "precision")


class Field(FieldBase):
    """0-D (constant), 1-D (z), 2-D (x, y) or 3-D gridded field."""

    def __init__(self, label="Field", dim=0, field=None, scaling=1.0, units="m", debug=False, method="linear",
                 interpolator_backend="auto", periodic_z=False, precision="fp64"):
        self._label = label
        self._dim = dim
        self._scaling = scaling
        self._debug = debug
        self._filename = None
        self._metadata = {}
        self._interpolator_backend = interpolator_backend
        self._method = method
        if units not in UNIT_SCALES:
            raise ValueError(f"Unknown unit '{units}'. Must be one of {list(UNIT_SCALES.keys())}")
        self._unit_scale = UNIT_SCALES[units]
        if precision not in _lib.PRECISIONS:
            raise ValueError(f"Unknown precision '{precision}'. Must be one of {list(_lib.PRECISIONS)}")
        self._precision = precision
        self._periodic_z = periodic_z
        self._is_zero = False
        self._field = {"x": None, "y": None, "z": None}
        if field is not None:
            self._field = field
        self._dev_map = None
        self._dev_spline = None

    # ---- core API ----------------------------------------------------------
    def __call__(self, pts):
        pts = np.atleast_2d(np.asarray(pts, dtype=np.float64))
        if pts.shape[1] != 3:
            raise ValueError(f"Points must have shape (M, 3), got {pts.shape}")
        m = pts.shape[0]
        if m == 0:
            return np.zeros((0, 3))
        pd = DeviceArray.from_host(pts)
        return evaluate_on_device(self, pd, m, self._precision).to_host()

    def __str__(self):
        if self._dim == 0:
            vals = [self._field[k] for k in ["x", "y", "z"]]
            return f"Field '{self._label}' (0D constant): {vals}"
        return f"Field '{self._label}' ({self._dim}D interpolated)"

    def __repr__(self):
        return f"Field(label='{self._label}', dim={self._dim}, scaling={self._scaling})"

    def plot(self, *args, **kwargs):
        raise NotImplementedError("plotting is outside the device engine")

    # ---- constructors ---------------------------------------------------------
    @classmethod
    def from_file(cls, filename, **kwargs):
        """field.py:284-343: load a map file, choosing the loader by extension.

        Keyword arguments that Field() takes (label, scaling, method, ...) go to the
        Field; the rest go to the loader (e.g. ``r_cutoff`` for .dat, ``mirror`` for
        .map).  Component arrays are reshaped onto the axes that have more than one
        node, so 2-D maps stored as (nx, ny, 1) load as 2-D Fields.  Pickled fields
        are not loaded: unpickling runs code from the file.
        """
        from .field_loaders import LOADERS
        if not os.path.exists(filename):
            raise FileNotFoundError(f"Field file not found: {filename}")
        _, ext = os.path.splitext(filename)
        if ext == ".pickle":
            raise NotImplementedError("pickled fields are not loaded (unpickling executes code from the file); "
                                      "save grid and values with numpy and use Field.from_arrays")
        if ext not in LOADERS:
            raise ValueError(f"Unknown file extension: {ext}")
        field_kw = {k: kwargs.pop(k) for k in list(kwargs) if k in _FIELD_KWARGS}
        data = LOADERS[ext](filename, **kwargs)
        if data.get("coordinate_system") == "cylindrical":
            raise ValueError("Field maps are Cartesian; load with cartesian_grid=True")
        field = cls(**field_kw)
        field._filename = filename
        pts = [np.asarray(data["grid"][k], dtype=np.float64) for k in ("x", "y", "z")
               if k in data["grid"] and len(data["grid"][k]) > 1]
        field._dim = data["dim"]
        field._metadata = data["metadata"]
        shape = tuple(len(p) for p in pts)
        for comp in ("x", "y", "z"):
            if comp in data["values"]:
                vals = np.asarray(data["values"][comp], dtype=np.float64)
                if vals.shape != shape and vals.size == int(np.prod(shape)):
                    vals = vals.reshape(shape)
                interp = get_interpolator(points=tuple(pts), values=vals, bounds_error=False, fill_value=0.0,
                                          method=field._method, backend=field._interpolator_backend)
                interp.periodic_z = field._periodic_z
                field._field[comp] = interp
        return field

    @classmethod
    def from_arrays(cls, grid, values, **kwargs):
        """field.py:345-387: one interpolator per component on the axes that have more than one node."""
        field = cls(**kwargs)
        grid_points = [np.asarray(grid[k], dtype=np.float64) for k in ["x", "y", "z"]
                       if k in grid and len(grid[k]) > 1]
        field._dim = len(grid_points)
        backend = field._interpolator_backend.lower()
        if backend not in BACKENDS:
            raise ValueError(f"Unknown backend '{backend}'. Use: 'auto', 'hip', 'numba', 'fast', 'scipy', 'cupy'")
        for component in ["x", "y", "z"]:
            if component in values:
                interp = get_interpolator(points=tuple(grid_points), values=values[component], bounds_error=False,
                                          fill_value=0.0, method=field._method, backend=backend)
                interp.periodic_z = field._periodic_z
                field._field[component] = interp
        return field

    @classmethod
    def zero(cls, dim=3):
        """field.py:389-407: zero everywhere."""
        if dim == 0:
            return cls(label="Zero Field", dim=0, field={"x": 0.0, "y": 0.0, "z": 0.0})
        f = cls(label="Zero Field", dim=dim)
        f._is_zero = True
        return f

    # ---- properties -----------------------------------------------------------
    @property
    def label(self):
        return self._label

    @label.setter
    def label(self, value):
        self._label = value

    @property
    def dim(self):
        return self._dim

    @property
    def scaling(self):
        return self._scaling

    @scaling.setter
    def scaling(self, value):
        self._scaling = value

    @property
    def metadata(self):
        return self._metadata

    @property
    def precision(self):
        return self._precision

    @precision.setter
    def precision(self, value):
        if value not in _lib.PRECISIONS:
            raise ValueError(f"Unknown precision '{value}'")
        self._precision = value

    # ---- device form ----------------------------------------------------------
    def _grid_source(self):
        """(points, [vx, vy, vz], periodic) of a gridded field, or None."""
        if self._is_zero or self._dim == 0:
            return None
        comps = self._shared_grid_components()
        if comps is None or any(c.order != 1 for c in comps):
            return None
        return comps[0].grid, [c.values for c in comps], self._periodic_z

    def _shared_grid_components(self):
        """The three HipInterpolator components when they share one grid, else None."""
        if self._is_zero or self._dim == 0:
            return None
        comps = [self._field.get(c) for c in "xyz"]
        if not all(isinstance(c, HipInterpolator) for c in comps):
            return None
        g0 = comps[0].grid
        for c in comps[1:]:
            if len(c.grid) != len(g0) or not all(np.array_equal(a, b) for a, b in zip(c.grid, g0)):
                return None
        return comps

    def device_spline(self):
        """One 3-component device spline for a field whose components share a grid and a spline
        order other than 1 (method 'nearest', 'cubic', 'quintic'), else None."""
        if self._dev_spline is None:
            comps = self._shared_grid_components()
            if comps is None or comps[0].order == 1 or any(c.order != comps[0].order or c.rgi != comps[0].rgi
                                                            or c.fill_value != comps[0].fill_value for c in comps):
                return None
            if self._periodic_z:
                raise NotImplementedError("periodic_z maps are interpolated linearly only")
            c0 = comps[0]
            self._dev_spline = DeviceSpline(c0.grid, [c.values for c in comps], c0.order, rgi=c0.rgi,
                                            cval=c0.fill_value)
        return self._dev_spline

    def device_map(self):
        if self._dev_map is None:
            src = self._grid_source()
            if src is None:
                raise ValueError("field has no gridded components")
            self._dev_map = DeviceFieldMap(src[0], src[1], 0.0, src[2])
        return self._dev_map

    def _term(self):
        """The device term of this field (PPField), or None if it needs per-component evaluation."""
        t = _lib.PPField()
        t.scale1 = 1.0
        if self._is_zero:
            t.kind = _lib.FIELD_ZERO
            return t
        if self._dim == 0:
            t.kind = _lib.FIELD_CONST
            for k, c in enumerate("xyz"):
                t.value[k] = self._scaling * self._field[c]   # field.py:451-453 (times ones(M))
            return t
        if self._grid_source() is None:
            return None
        t.kind = _lib.FIELD_MAP
        t.map = self.device_map().handle
        t.comp0 = 0
        t.scale1 = float(self._scaling)
        return t


class CompositeField(FieldBase):
    """Linear combination sum_i w_i F_i (field.py:595-658)."""

    def __init__(self, fields, weights):
        if len(fields) != len(weights):
            raise ValueError("Number of fields must match number of weights")
        self.fields = fields
        self.weights = np.array(weights, dtype=np.float64)
        self.n_fields = len(fields)

    def __call__(self, pts):
        pts = np.atleast_2d(np.asarray(pts, dtype=np.float64))
        m = len(pts)
        if m == 0:
            return np.zeros((0, 3))
        pd = DeviceArray.from_host(pts)
        return evaluate_on_device(self, pd, m, _precision_of(self)).to_host()

    def __str__(self):
        return f"CompositeField with {len(self.fields)} components"


class ScaledField(FieldBase):
    """scale * F + offset (field.py:661-690)."""

    def __init__(self, field, scale=1.0, offset=0.0):
        self.field = field
        self.scale = scale
        self.offset = offset

    def __call__(self, pts):
        pts = np.atleast_2d(np.asarray(pts, dtype=np.float64))
        m = len(pts)
        if m == 0:
            return np.zeros((0, 3))
        pd = DeviceArray.from_host(pts)
        return evaluate_on_device(self, pd, m, _precision_of(self)).to_host()


class RFField(FieldBase):
    """Time-dependent field F(x) * cos(omega t + phase) (mode 'cos') or sin (extension).

    Called with points only it returns F(x) (the static profile); the pushers
    evaluate it at t_n = t0 + n dt and at the RK4 stage times
    t_n + (0, 1/2, 1/2, 1) dt.
    """

    time_dependent = True

    def __init__(self, field, omega, phase=0.0, mode="cos"):
        if mode not in ("cos", "sin"):
            raise ValueError("mode must be 'cos' or 'sin'")
        self.field = field
        self.omega = float(omega)
        self.phase = float(phase)
        self.mode = mode

    def factor(self, t):
        arg = self.omega * t + self.phase
        return np.cos(arg) if self.mode == "cos" else np.sin(arg)

    def __call__(self, pts, t=None):
        pts = np.atleast_2d(np.asarray(pts, dtype=np.float64))
        m = len(pts)
        if m == 0:
            return np.zeros((0, 3))
        pd = DeviceArray.from_host(pts)
        return evaluate_on_device(self, pd, m, _precision_of(self), t).to_host()


class FieldMap:
    """Container of an electric and a magnetic field (field.py:697-755)."""

    def __init__(self, efield=None, bfield=None, metadata=None):
        self.efield = efield if efield is not None else Field.zero()
        self.bfield = bfield if bfield is not None else Field.zero()
        self.metadata = metadata if metadata is not None else {}

    @classmethod
    def from_h5part(cls, filename, iteration=0):
        raise NotImplementedError("H5Part FieldMap loader is not part of the device engine")

    @classmethod
    def from_fields(cls, efield, bfield, **metadata):
        return cls(efield=efield, bfield=bfield, metadata=metadata)

    def __str__(self):
        return f"FieldMap(E: {self.efield}, B: {self.bfield})"


# ---------------------------------------------------------------------------
# Compilation of field expressions
# ---------------------------------------------------------------------------

def _precision_of(f):
    if isinstance(f, Field):
        return f._precision
    for attr in ("field",):
        if hasattr(f, attr):
            return _precision_of(getattr(f, attr))
    if isinstance(f, CompositeField) and f.fields:
        return _precision_of(f.fields[0])
    return "fp64"


def compile_term(f):
    """A PPField the push kernels read directly, or None.

    Field                      -> CONST / MAP term (scale1 = Field.scaling)
    ScaledField(Field)         -> same term with scale2 / offset
    RFField(Field or Scaled)   -> same term with the cos/sin column of the rf table
    """
    if isinstance(f, Field):
        return f._term()
    if isinstance(f, ScaledField) and isinstance(f.field, Field):
        t = f.field._term()
        if t is None:
            return None
        t.scale2 = float(f.scale)
        t.has_scale2 = 1
        t.offset = float(f.offset)
        t.has_offset = 1
        return t
    if isinstance(f, RFField):
        t = compile_term(f.field)
        if t is None or t.rf != _lib.RF_NONE:
            return None
        t.rf = _lib.RF_COS if f.mode == "cos" else _lib.RF_SIN
        return t
    return None


def is_time_dependent(f):
    if isinstance(f, RFField):
        return True
    if isinstance(f, ScaledField):
        return is_time_dependent(f.field)
    if isinstance(f, CompositeField):
        return any(is_time_dependent(g) for g in f.fields)
    return bool(getattr(f, "time_dependent", False))


def evaluate_on_device(f, pts_dev, m, precision, t=None):
    """Evaluate any field expression at device points (m, 3) into a new DeviceArray.

    Combination steps follow the reference's NumPy order: ScaledField is
    ``scale * F + offset``, CompositeField accumulates ``total += w * F`` from
    zeros.  Arbitrary Python callables are evaluated on the host (their own
    code) and uploaded.
    """
    out = DeviceArray((m, 3))
    if isinstance(f, RFField):
        inner = evaluate_on_device(f.field, pts_dev, m, precision, None)
        if t is None:
            return inner
        _lib.call("pp_scale_offset_batch", inner.ptr, float(f.factor(t)), 1, 0.0, 0, m, out.ptr, None)
        return out
    term = compile_term(f)
    if term is not None:
        return eval_term(term, pts_dev, m, precision, 1.0, out)
    if isinstance(f, ScaledField):
        inner = evaluate_on_device(f.field, pts_dev, m, precision, t)
        _lib.call("pp_scale_offset_batch", inner.ptr, float(f.scale), 1, float(f.offset), 1, m, out.ptr, None)
        return out
    if isinstance(f, CompositeField):
        out.fill(0.0)
        for g, w in zip(f.fields, f.weights):
            val = evaluate_on_device(g, pts_dev, m, precision, t)
            _lib.call("pp_axpy_batch", out.ptr, float(w), None, val.ptr, m, out.ptr, None)
        return out
    if isinstance(f, Field) and f.device_spline() is not None:
        # spline methods: the three components from device points, 1-D fields on z, 2-D on x, y
        f.device_spline().eval_device(pts_dev, m, out, stride=3, col0=2 if f._dim == 1 else 0)
        _lib.call("pp_scale_offset_batch", out.ptr, float(f._scaling), 1, 0.0, 0, m, out.ptr, None)
        return out
    if isinstance(f, Field):
        # components on different grids (or custom callables): one device gather per component
        host_pts = pts_dev.to_host()
        q = host_pts[:, 2:3] if f._dim == 1 else (host_pts[:, :2] if f._dim == 2 else host_pts)
        cols = []
        for c in "xyz":
            interp = f._field[c]
            cols.append(interp(q, precision=precision) if isinstance(interp, HipInterpolator) else interp(q))
        out.upload(np.column_stack(cols))
        _lib.call("pp_scale_offset_batch", out.ptr, float(f._scaling), 1, 0.0, 0, m, out.ptr, None)
        return out
    # arbitrary callable: the user's own code runs on the host
    host_pts = pts_dev.to_host()
    vals = f(host_pts, t) if getattr(f, "time_dependent", False) else f(host_pts)
    vals = np.asarray(vals, dtype=np.float64)
    if vals.shape != (m, 3):
        raise ValueError(f"field callable returned shape {vals.shape}, expected ({m}, 3)")
    out.upload(vals)
    return out


# ---------------------------------------------------------------------------
# Coordinate helpers (field.py:762-824)
# ---------------------------------------------------------------------------

def cartesian_to_cylindrical(x, y, z):
    return np.sqrt(x ** 2 + y ** 2), np.arctan2(y, x), z


def cylindrical_to_cartesian(r, theta, z):
    return r * np.cos(theta), r * np.sin(theta), z


def transform_cylindrical_field_to_cartesian(r, theta, z, fr, ftheta, fz):
    x = r * np.cos(theta)
    y = r * np.sin(theta)
    fx = fr * np.cos(theta) - ftheta * np.sin(theta)
    fy = fr * np.sin(theta) + ftheta * np.cos(theta)
    return x, y, z, fx, fy, fz
