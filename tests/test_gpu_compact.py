"""Alive-mask compaction (pp_compact_state) and the fused loss test (pp_push_lossy).

Compaction is a stable partition: live slots first, dead slots after, each part in
(map cell, original index) order; ids, the six state arrays and the alive mask move
together, and pp_unsort_state restores the original order bit for bit.
"""
import numpy as np
import pytest

import pypatools_amd as pa
from pypatools_amd.device import DeviceArray
from pypatools_amd.engine import FieldPair, ParticleState

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _device():
    from pypatools_amd import device
    if device.device_count() == 0:
        pytest.skip("no HIP device visible")


def _grid(n=12, dx=1e-3):
    x = (np.arange(n) - (n - 1) / 2.0) * dx
    z = np.arange(n) * dx
    return {"x": x, "y": x, "z": z}


def _map_field(grid):
    shape = (len(grid["x"]), len(grid["y"]), len(grid["z"]))
    rng = np.random.default_rng(3)
    return pa.Field.from_arrays(grid, {c: rng.standard_normal(shape) for c in "xyz"})


def _cell_keys(grid, r):
    """The coefficient-table cell of each position: floor((x - lo) * inv_d), clamped to the cells."""
    keys = np.zeros(len(r), dtype=np.int64)
    for a, c in enumerate("xyz"):
        g = grid[c]
        nc = len(g) - 1
        f = np.clip(np.floor((r[:, a] - g[0]) * ((len(g) - 1) / (g[-1] - g[0]))), 0, nc - 1).astype(np.int64)
        keys = keys * nc + f
    return keys


def _state(r, v, alive):
    st = ParticleState.from_host(r, v)
    st.alive = DeviceArray.from_host(alive.astype(np.uint8))
    return st


@pytest.mark.parametrize("with_map", [True, False])
@pytest.mark.parametrize("n", [1, 777, 20000])
def test_compaction_is_a_stable_partition_and_unsort_restores(with_map, n):
    rng = np.random.default_rng(n + with_map)
    grid = _grid()
    r = np.column_stack([rng.uniform(-7e-3, 7e-3, n), rng.uniform(-7e-3, 7e-3, n), rng.uniform(-1e-3, 13e-3, n)])
    v = rng.standard_normal((n, 3)) * 1e5
    alive = rng.uniform(size=n) < 0.6
    fp = FieldPair(_map_field(grid), pa.Field.zero()) if with_map else FieldPair(pa.Field.zero(), pa.Field.zero())
    st = _state(r, v, alive)
    n_live = st.compact(fp)
    assert n_live == int(alive.sum())
    ids = st.ids.to_host()
    key = _cell_keys(grid, r) if with_map else np.zeros(n, dtype=np.int64)
    want = np.lexsort((np.arange(n), key, ~alive))          # dead last, then cell, then original index
    np.testing.assert_array_equal(ids, want)
    rs, vs = st.to_host()
    np.testing.assert_array_equal(rs, r[want])
    np.testing.assert_array_equal(vs, v[want])
    np.testing.assert_array_equal(st.alive.to_host(), (np.arange(n) < n_live).astype(np.uint8))
    st.unsort()
    rs, vs = st.to_host()
    np.testing.assert_array_equal(rs, r)
    np.testing.assert_array_equal(vs, v)
    np.testing.assert_array_equal(st.alive.to_host(), alive.astype(np.uint8))


def test_repeated_compaction_follows_new_losses():
    rng = np.random.default_rng(9)
    n = 5000
    grid = _grid()
    r = np.column_stack([rng.uniform(-5e-3, 5e-3, n), rng.uniform(-5e-3, 5e-3, n), rng.uniform(0, 11e-3, n)])
    v = rng.standard_normal((n, 3))
    alive = np.ones(n, dtype=bool)
    fp = FieldPair(_map_field(grid), pa.Field.zero())
    st = _state(r, v, alive)
    for rnd in range(4):
        n_live = st.compact(fp)
        assert n_live == int(alive.sum())
        ids = st.ids.to_host()
        # kill some live slots in slot order, as a lossy push would
        kill = rng.uniform(size=n_live) < 0.3
        mask = st.alive.to_host()
        mask[:n_live][kill] = 0
        st.alive.upload(mask)
        alive[ids[:n_live][kill]] = False
    st.unsort()
    np.testing.assert_array_equal(st.alive.to_host(), alive.astype(np.uint8))
    rs, _ = st.to_host()
    np.testing.assert_array_equal(rs, r)


def test_all_dead_and_all_alive():
    rng = np.random.default_rng(4)
    n = 300
    r, v = rng.standard_normal((n, 3)) * 1e-3, rng.standard_normal((n, 3))
    fp = FieldPair(pa.Field.zero(), pa.Field.zero())
    st = _state(r, v, np.zeros(n, dtype=bool))
    assert st.compact(fp) == 0
    np.testing.assert_array_equal(st.ids.to_host(), np.arange(n))
    st = _state(r, v, np.ones(n, dtype=bool))
    assert st.compact(fp) == n
    np.testing.assert_array_equal(st.ids.to_host(), np.arange(n))


def test_compaction_and_lossy_push_reject_bad_arguments():
    import ctypes

    from pypatools_amd import _lib
    from pypatools_amd.electrodes import TriangleElectrodes
    st = ParticleState.from_host(np.zeros((10, 3)), np.zeros((10, 3)))
    fp = FieldPair(pa.Field.zero(), pa.Field.zero())
    ws = st._workspace()
    a = st._state_args(fp)                                   # no alive mask
    with pytest.raises(Exception, match="alive"):
        _lib.call("pp_compact_state", a, None, 2, 0, st.ensure_ids().ptr, ctypes.byref(ctypes.c_int64()),
                  ws.ptr, ws.nbytes, None)
    mesh = TriangleElectrodes.box([-1, -1, -1], [1, 1, 1])
    a.algorithm, a.precision, a.nsteps, a.dt = _lib.ALG_IDS["boris"], _lib.PRECISIONS["fp64"], 1, 1e-9
    with pytest.raises(Exception, match="alive"):
        _lib.call("pp_push_lossy", a, mesh.handle, None)
    alive = DeviceArray.from_host(np.ones(10, dtype=np.uint8))
    a.alive = alive.ptr
    a.hist_rows = 5
    with pytest.raises(Exception, match="hist_rows"):
        _lib.call("pp_push_lossy", a, mesh.handle, None)
    with pytest.raises(Exception, match="mesh"):
        _lib.call("pp_push_lossy", a, None, None)
