"""Pins of the CoordinateMapper restatement of every spline order (oracle/interp.py), beyond its being
scipy.ndimage.map_coordinates itself: node values are reproduced (the prefilter makes the spline
interpolate), order 0 rounds half-way coordinates up, order 1 is exact on linear data, points outside
the grid get the fill value; and the 'scipy' backend's 'nearest' sends a point at a cell's midpoint
to the lower node."""
import numpy as np
import pytest

from oracle.interp import interp_coordinate_mapper, interp_rgi_nearest


def _grid(rng, n, nonuniform):
    g = np.cumsum(rng.uniform(0.5, 1.5, n)) if nonuniform else np.arange(n, dtype=float)
    return g - g[0] + rng.normal()


@pytest.mark.parametrize("order", [0, 1, 3, 5])
@pytest.mark.parametrize("nonuniform", [False, True])
def test_spline_reproduces_node_values(order, nonuniform):
    rng = np.random.default_rng(order + 10 * nonuniform)
    pts = (_grid(rng, 7, nonuniform), _grid(rng, 5, nonuniform), _grid(rng, 9, nonuniform))
    vals = rng.normal(size=(7, 5, 9))
    I, J, K = np.meshgrid(np.arange(7), np.arange(5), np.arange(9), indexing="ij")
    xi = np.column_stack([pts[0][I.ravel()], pts[1][J.ravel()], pts[2][K.ravel()]])
    for be in (False, True):
        out = interp_coordinate_mapper(pts, vals, xi, 0.0, order=order, bounds_error=be)
        np.testing.assert_allclose(out, vals.ravel(), rtol=0, atol=1e-12)


def test_order0_rounds_half_up_and_order1_is_exact_on_linear_data():
    g = np.arange(6.0)
    v = 3.0 * g - 1.0
    xi = np.array([[0.5], [1.5], [2.5], [4.49], [4.5], [5.0]])
    np.testing.assert_array_equal(interp_coordinate_mapper((g,), v, xi, 0.0, order=0), v[[1, 2, 3, 4, 5, 5]])
    x = np.linspace(0, 5, 41)[:, None]
    np.testing.assert_allclose(interp_coordinate_mapper((g,), v, x, 0.0, order=1), 3.0 * x[:, 0] - 1.0, atol=1e-14)


@pytest.mark.parametrize("order", [0, 1, 3, 5])
def test_outside_gets_fill(order):
    g = np.arange(5.0)
    v = np.arange(5.0) ** 2
    xi = np.array([[-0.01], [4.01], [-3.0], [9.0]])
    np.testing.assert_array_equal(interp_coordinate_mapper((g,), v, xi, -7.0, order=order), np.full(4, -7.0))


def test_rgi_nearest_ties_go_to_the_lower_node():
    g = np.array([0.0, 1.0, 3.0])
    v = np.array([10.0, 20.0, 30.0])
    xi = np.array([[0.5], [0.51], [2.0], [2.0001], [3.0], [3.5]])
    np.testing.assert_array_equal(interp_rgi_nearest((g,), v, xi, -1.0), [10.0, 20.0, 20.0, 30.0, 30.0, -1.0])
