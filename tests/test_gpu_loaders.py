"""A map loaded from an OPERA table drives the device push exactly like the same arrays given to from_arrays."""
import os

import numpy as np
import pytest

import pypatools_amd as pa
from oracle.field import OField
from oracle.pusher import OraclePusher
from pypatools_amd.workloads import c2_map

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _device():
    from pypatools_amd import device
    if device.device_count() == 0:
        pytest.skip("no HIP device visible")


def _table(path, grid, vals, kind):
    xs, ys, zs = (grid[a] * 100.0 for a in "xyz")                      # cm in the file
    X, Y, Z = np.meshgrid(xs, ys, zs, indexing="ij")
    unit = "[Tesla]" if kind == "B" else "[V/m]"
    head = [f"{len(xs)} {len(ys)} {len(zs)} 2", "1 X [cm]", "2 Y [cm]", "3 Z [cm]",
            f"4 {kind}X {unit}", f"5 {kind}Y {unit}", f"6 {kind}Z {unit}", "0"]
    data = np.column_stack([X.ravel(), Y.ravel(), Z.ravel()] + [vals[c].ravel() for c in "xyz"])
    with open(path, "w") as f:
        f.write("\n".join(head) + "\n")
        np.savetxt(f, data, fmt="%.17g")


def test_from_file_track_equals_from_arrays_and_oracle(tmp_path):
    grid, E, B = c2_map(12)
    pe, pb = os.path.join(str(tmp_path), "e.table"), os.path.join(str(tmp_path), "b.table")
    _table(pe, grid, E, "E")
    _table(pb, grid, B, "B")
    fe, fb = pa.Field.from_file(pe), pa.Field.from_file(pb)
    lg = {a: np.unique(np.asarray(fe._field["x"].grid[i])) for i, a in enumerate("xyz")}
    le = {c: fe._field[c].values for c in "xyz"}
    lb = {c: fb._field[c].values for c in "xyz"}
    rng = np.random.default_rng(41)
    n = 300
    r0 = np.column_stack([rng.normal(0, 1e-3, n), rng.normal(0, 1e-3, n), rng.uniform(1e-3, 4e-3, n)])
    v0 = np.column_stack([rng.normal(0, 1e4, n), rng.normal(0, 1e4, n), np.full(n, 1.5e6)])
    qm = pa.IonSpecies("proton").q_over_m
    p = pa.Pusher(pa.IonSpecies("proton"), "boris")
    rh, vh, _ = p.track_batch(r0, v0, fe, fb, 120, 5e-11, 40)
    ga, gb = pa.Field.from_arrays(lg, le), pa.Field.from_arrays(lg, lb)
    rh2, vh2, _ = p.track_batch(r0, v0, ga, gb, 120, 5e-11, 40)
    np.testing.assert_array_equal(rh, rh2)
    np.testing.assert_array_equal(vh, vh2)
    orc = OraclePusher(qm, "boris")
    orh, ovh, _ = orc.track_batch(r0, v0, OField.from_arrays(lg, le), OField.from_arrays(lg, lb), 120, 5e-11, 40)
    np.testing.assert_array_equal(rh, orh)
    np.testing.assert_array_equal(vh, ovh)
