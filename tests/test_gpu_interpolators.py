"""get_interpolator(...) callables and spline Fields on the device against the oracle's
CoordinateMapper / RegularGridInterpolator restatement (interpolators.py:415-610) and the
reference's own 'fast' output (golden).

Tolerances: order 0 and the 'scipy' backend's 'nearest' pick node values (bit for bit); order 1
agrees to 4e-15 of the value range (the device gather is NumbaInterpolator's nested-lerp form of
the same bilinear/trilinear weights: up to seven lerps on non-uniform 3-D grids, a few ulps);
orders 3 and 5 to 1e-12 of the value range (the prefilter recursions and the tap sums round in a
different order than scipy's)."""
import numpy as np
import pytest

import pypatools_amd as pa
from oracle.field import OField
from oracle.interp import interp_coordinate_mapper, interp_rgi_nearest
from oracle.pusher import OraclePusher

pytestmark = pytest.mark.gpu

TOL = {0: 0.0, 1: 4e-15, 3: 1e-12, 5: 1e-12}
METHODS = {"nearest": 0, "linear": 1, "cubic": 3, "quintic": 5}


@pytest.fixture(scope="module", autouse=True)
def _device():
    from pypatools_amd import device
    if device.device_count() == 0:
        pytest.skip("no HIP device visible")


def _grid(rng, n, nonuniform):
    g = np.cumsum(rng.uniform(0.5, 1.5, n)) if nonuniform else np.arange(n, dtype=float)
    return (g - g[0]) * 1e-3 + rng.normal() * 1e-2


def _points(rng, pts, m):
    """Points inside, on nodes, on faces and outside the grid."""
    lo = np.array([p[0] for p in pts])
    hi = np.array([p[-1] for p in pts])
    span = hi - lo
    xi = lo + rng.uniform(-0.1, 1.1, (m, len(pts))) * span
    k = m // 4
    for a, p in enumerate(pts):
        xi[:k, a] = rng.choice(p, k)                         # on nodes
    xi[k:k + 8] = hi
    xi[k + 8:k + 16] = lo
    return xi


def _check(out, ref, vals, order):
    scale = max(np.max(np.abs(vals)), 1e-300)
    err = np.max(np.abs(out - ref)) / scale if len(out) else 0.0
    assert err <= TOL[order], (order, err)


@pytest.mark.parametrize("ndim", [1, 2, 3])
@pytest.mark.parametrize("nonuniform", [False, True])
@pytest.mark.parametrize("method", list(METHODS))
def test_coordinate_mapper_semantics(ndim, nonuniform, method):
    rng = np.random.default_rng(100 * ndim + 10 * nonuniform + METHODS[method])
    shape = (9, 6, 11)[:ndim]
    pts = tuple(_grid(rng, n, nonuniform) for n in shape)
    vals = rng.normal(size=shape) * 3.0
    xi = _points(rng, pts, 3000)
    order = METHODS[method]
    for backend in ("auto", "fast", "hip"):
        it = pa.get_interpolator(pts, vals, method=method, backend=backend, fill_value=-2.5)
        out = it(xi)
        ref = interp_coordinate_mapper(pts, vals, xi, -2.5, order=order)
        _check(out, ref, vals, order)
    # single point, and the per-call override (interpolators.py:482-485)
    it = pa.get_interpolator(pts, vals, method=method)
    p0 = xi[len(xi) // 2]
    assert isinstance(it(p0), float)
    _check(np.array([it(p0)]), interp_coordinate_mapper(pts, vals, p0[None], 0.0, order=order), vals, order)
    for call in ("nearest", "linear", "cubic", "quintic"):
        o = {"nearest": 0, "linear": 1, "cubic": 3}.get(call, order)
        _check(it(xi, method=call), interp_coordinate_mapper(pts, vals, xi, 0.0, order=o), vals, o)


@pytest.mark.parametrize("method", list(METHODS))
def test_bounds_error_mode_nearest(method):
    rng = np.random.default_rng(7 + METHODS[method])
    pts = (_grid(rng, 8, True), _grid(rng, 7, False), _grid(rng, 10, True))
    vals = rng.normal(size=(8, 7, 10))
    it = pa.get_interpolator(pts, vals, method=method, bounds_error=True)
    xi = np.column_stack([rng.uniform(p[0], p[-1], 2000) for p in pts])
    xi[:6] = [p[0] for p in pts]
    xi[6:12] = [p[-1] for p in pts]
    order = METHODS[method]
    _check(it(xi), interp_coordinate_mapper(pts, vals, xi, 0.0, order=order, bounds_error=True), vals, order)
    with pytest.raises(ValueError, match="outside the interpolation domain"):
        it(np.array([[pts[0][0] - 1e-6, pts[1][1], pts[2][1]]]))


def test_scipy_backend_nearest_and_linear():
    rng = np.random.default_rng(3)
    pts = (_grid(rng, 9, True), _grid(rng, 8, True))
    vals = rng.normal(size=(9, 8))
    xi = _points(rng, pts, 4000)
    # points exactly at cell midpoints: RegularGridInterpolator sends them to the lower node
    xi[-50:, 0] = 0.5 * (pts[0][2] + pts[0][3])
    it = pa.get_interpolator(pts, vals, method="nearest", backend="scipy", fill_value=9.0)
    np.testing.assert_array_equal(it(xi), interp_rgi_nearest(pts, vals, xi, 9.0))
    lin = pa.get_interpolator(pts, vals, method="linear", backend="scipy", fill_value=9.0)
    _check(lin(xi), interp_coordinate_mapper(pts, vals, xi, 9.0, order=1), vals, 1)
    np.testing.assert_array_equal(lin(xi, method="nearest"), interp_rgi_nearest(pts, vals, xi, 9.0))


def test_fast_backend_matches_reference_golden(golden_dir):
    """The reference's own 'fast' (CoordinateMapper, linear) output for Field.__call__ (field_call.npz)."""
    import os
    g = np.load(os.path.join(golden_dir, "field_call.npz"))
    grid = {"x": g["gx"], "y": g["gy"], "z": g["gz"]}
    pts3 = (g["gx"], g["gy"], g["gz"])
    for fb, scaling in (("E", 1.25), ("B", 1.0)):
        vals = {c: g[f"val_{fb}{c}"] for c in "xyz"}
        ref = g[f"{fb}3d_fast"]
        f = pa.Field.from_arrays(grid, vals, scaling=scaling, interpolator_backend="fast")
        np.testing.assert_allclose(f(g["pts"]), ref, rtol=0, atol=1e-15 * np.max(np.abs(ref)))
        # the component interpolators called directly, single point and batch
        for k, c in enumerate("xyz"):
            it = pa.get_interpolator(pts3, vals[c], backend="fast")
            np.testing.assert_allclose(scaling * it(g["pts"]), ref[:, k], rtol=0, atol=1e-15 * np.max(np.abs(ref)))
            assert abs(scaling * it(g["pts"][7]) - ref[7, k]) <= 1e-15 * np.max(np.abs(ref))


@pytest.mark.parametrize("method", ["nearest", "cubic", "quintic"])
def test_spline_field_and_track_against_oracle(method):
    """Field(method=...) on the device, and a track through it against the oracle pusher."""
    from pypatools_amd.workloads import c2_map
    grid, E, B = c2_map(24)
    order = METHODS[method]
    fe = pa.Field.from_arrays(grid, E, method=method)
    fb = pa.Field.from_arrays(grid, B, method=method, scaling=0.8)
    oe = OField.from_arrays(grid, E, interp="coordmapper", order=order)
    ob = OField.from_arrays(grid, B, scaling=0.8, interp="coordmapper", order=order)
    rng = np.random.default_rng(5)
    pts = np.column_stack([rng.normal(0, 4e-3, 3000), rng.normal(0, 4e-3, 3000), rng.uniform(-1e-3, 0.025, 3000)])
    scale = max(np.max(np.abs(oe(pts))), 1e-300)
    assert np.max(np.abs(fe(pts) - oe(pts))) / scale <= TOL[order]
    scale = max(np.max(np.abs(ob(pts))), 1e-300)
    assert np.max(np.abs(fb(pts) - ob(pts))) / scale <= TOL[order]
    qm = pa.IonSpecies("proton").q_over_m
    M, N, dt = 256, 60, 5e-11
    r0 = np.column_stack([rng.normal(0, 2e-3, M), rng.normal(0, 2e-3, M), rng.uniform(2e-3, 6e-3, M)])
    v0 = np.column_stack([rng.normal(0, 1e4, M), rng.normal(0, 1e4, M), np.full(M, 2.19e7)])
    orh, ovh, _ = OraclePusher(qm, "boris").track_batch(r0, v0, oe, ob, N, dt, 20)
    rh, vh, _ = pa.Pusher(pa.IonSpecies("proton"), "boris").track_batch(r0, v0, fe, fb, N, dt, 20)
    path = np.linalg.norm(v0, axis=1) * N * dt
    err = np.max(np.linalg.norm(rh[-1] - orh[-1], axis=1) / path)
    assert err < 1e-11, err
