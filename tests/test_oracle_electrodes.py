"""The brute-force segment / mesh oracle against geometry that fixes the answer."""
import numpy as np

from oracle.electrodes import segment_hits

PLATE_V = np.array([[-1.0, -1.0, 0.0], [1.0, -1.0, 0.0], [1.0, 1.0, 0.0], [-1.0, 1.0, 0.0]])
PLATE_T = np.array([[0, 1, 2], [0, 2, 3]])


def _box():
    lo, hi = np.array([-1.0, -2.0, 0.5]), np.array([1.5, 2.0, 3.0])
    v = np.array([[lo[0] if i & 1 == 0 else hi[0], lo[1] if i & 2 == 0 else hi[1], lo[2] if i & 4 == 0 else hi[2]]
                  for i in range(8)])
    quads = [(0, 1, 3, 2), (4, 6, 7, 5), (0, 4, 5, 1), (2, 3, 7, 6), (0, 2, 6, 4), (1, 5, 7, 3)]
    return lo, hi, v, np.array([t for a, b, c, d in quads for t in ((a, b, c), (a, c, d))])


def test_plate_closed_forms():
    r0 = np.array([[0.2, 0.3, -1.0], [0.2, 0.3, 1.0], [2.0, 0.0, -1.0], [0.0, 0.0, -1.0], [0.5, 0.5, 0.0],
                   [-2.0, 0.1, 0.0], [1.0, 1.0, -1.0], [0.3, 0.3, 1e-12]])
    r1 = np.array([[0.2, 0.3, 1.0], [0.2, 0.3, 2.0], [2.0, 0.0, 1.0], [0.0, 0.0, 0.0], [0.5, 0.5, 1.0],
                   [2.0, 0.1, 0.0], [1.0, 1.0, 1.0], [0.3, 0.3, 5e-12]])
    want = [True,     # crosses inside the square
            False,    # stays above
            False,    # crosses the plane outside the square
            True,     # ends on the plate (t = 1)
            True,     # starts on the plate (t = 0)
            False,    # lies in the plate's plane: det = 0
            True,     # through the corner vertex (closed edges)
            False]    # stops short of the plane
    np.testing.assert_array_equal(segment_hits(PLATE_V, PLATE_T, r0, r1), want)


def test_closed_box_one_endpoint_inside_always_hits():
    lo, hi, v, t = _box()
    rng = np.random.default_rng(9)
    inside = lo + (hi - lo) * rng.uniform(0.01, 0.99, (400, 3))
    outside = lo + (hi - lo) * rng.uniform(-2.0, 3.0, (400, 3))
    out_mask = np.any((outside < lo) | (outside > hi), axis=1)
    h = segment_hits(v, t, inside, outside)
    assert np.all(h[out_mask])                              # leaving a closed surface crosses it
    inside2 = lo + (hi - lo) * rng.uniform(0.01, 0.99, (400, 3))
    assert not np.any(segment_hits(v, t, inside, inside2))  # convex: two inside points never cross
    assert out_mask.sum() > 300
