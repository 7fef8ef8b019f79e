"""Pin the CPU oracle against golden vectors written by the reference itself.

The vectors come from tests/golden/make_golden.py, which ran PyPATools in this
container (no numba, so its NumPy paths).  The oracle's *as-shipped* form must
reproduce them bit for bit; its *canonical* form (``x*x`` squares, Numba-style
gather) must agree to the stated tolerances.
"""
import os

import numpy as np
import pytest

from oracle import kernels as K
from oracle import species as OS
from oracle.field import OComposite, OField, OScaled, OZero
from oracle.pusher import REFERENCE_ALGORITHMS, OraclePusher


def load(golden_dir, name):
    return np.load(os.path.join(golden_dir, name), allow_pickle=False)


def rel_err(a, b):
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    assert np.array_equal(np.isnan(a), np.isnan(b))
    m = ~np.isnan(a)
    scale = max(np.max(np.abs(b[m])), 1e-300) if np.any(m) else 1.0
    return float(np.max(np.abs(a[m] - b[m])) / scale) if np.any(m) else 0.0


def bitwise(a, b):
    np.testing.assert_array_equal(np.asarray(a), np.asarray(b))


def test_species_q_over_m_bitwise(golden_dir):
    g = load(golden_dir, "species_qm.npz")
    for i, name in enumerate(g["names"]):
        name = str(name)
        assert OS.q_over_m(name) == g["q_over_m"][i], name
        assert OS.mass_mev(name) == g["mass_mev"][i], name
        assert OS.mass_kg(name) == g["mass_kg"][i], name


def test_vay_rel_kernel_as_shipped_bitwise(golden_dir):
    g = load(golden_dir, "vay_single.npz")
    for i in range(len(g["dt"])):
        out = K.vay_rel_kick(g["v"][i:i + 1], g["E"][i:i + 1], g["B"][i:i + 1], g["dt"][i], g["qm"][i], K.sq_pow)
        bitwise(out[0], g["out"][i])


def test_vay_rel_kernel_canonical_close(golden_dir):
    g = load(golden_dir, "vay_single.npz")
    n = len(g["dt"])
    out = np.vstack([K.vay_rel_kick(g["v"][i:i + 1], g["E"][i:i + 1], g["B"][i:i + 1], g["dt"][i], g["qm"][i])
                     for i in range(n)])
    err = np.max(np.abs(out - g["out"]) / np.maximum(np.linalg.norm(g["out"], axis=1), 1e-300)[:, None])
    assert err < 1e-13


def _map_fields(g, backend_interp, pref="val_"):
    grid = {"x": g["gx"], "y": g["gy"], "z": g["gz"]}
    E = {c: g[f"{pref}E{c}"] for c in "xyz"}
    B = {c: g[f"{pref}B{c}"] for c in "xyz"}
    return grid, E, B


@pytest.mark.parametrize("interp,exact", [("coordmapper", True), ("numba", False)])
def test_field_call_3d(golden_dir, interp, exact):
    g = load(golden_dir, "field_call.npz")
    grid, E, B = _map_fields(g, interp)
    fE = OField.from_arrays(grid, E, scaling=1.25, interp=interp)
    fB = OField.from_arrays(grid, B, interp=interp)
    pts = g["pts"]
    for ours, key in ((fE(pts), "E3d_fast"), (fB(pts), "B3d_fast")):
        if exact:
            bitwise(ours, g[key])
        else:
            assert rel_err(ours, g[key]) < 1e-15
    # the scipy RegularGridInterpolator backend computes the same trilinear form
    assert rel_err(fE(pts), g["E3d_scipy"]) < 1e-15
    assert rel_err(fB(pts), g["B3d_scipy"]) < 1e-15


def test_field_call_1d_2d_0d_algebra(golden_dir):
    g = load(golden_dir, "field_call.npz")
    pts = g["pts"]
    assert str(g["fast1d_error"]) == "ValueError"  # reference quirk, see make_golden.py
    g1 = {"x": np.array([0.0]), "y": np.array([0.0]), "z": g["gz"]}
    f1 = OField.from_arrays(g1, {c: g["val1d"][i] for i, c in enumerate("xyz")}, scaling=0.5)
    assert rel_err(f1(pts), g["B1d"]) < 1e-15
    g2 = {"x": g["gx"], "y": g["gy"], "z": np.array([0.0])}
    v2 = {c: g["val2d"][i] for i, c in enumerate("xyz")}
    bitwise(OField.from_arrays(g2, v2, interp="coordmapper")(pts), g["E2d"])
    assert rel_err(OField.from_arrays(g2, v2)(pts), g["E2d"]) < 1e-15
    c0 = OField.constant(0.3, -1.1, 2.5, scaling=3.0)
    bitwise(c0(pts), g["const0d"])
    grid, _E, B = _map_fields(g, "coordmapper")
    fB = OField.from_arrays(grid, B, interp="coordmapper")
    bitwise(OScaled(fB, scale=2.5)(pts), g["scaled"])
    bitwise(OScaled(fB, offset=0.125)(pts), g["offset"])
    bitwise(OComposite([fB, c0], [1.0, -1.0])(pts), g["composite"])
    bitwise(OZero()(pts), g["zero3"])


def _uniform_fields(g):
    Ec, Bc = g["Ec"], g["Bc"]
    return (OField.constant(*Ec, scaling=float(g["E_scaling"])), OField.constant(*Bc))


# canonical-vs-reference tolerance (relative to the largest |value| of the array):
# the two differ only by libm pow vs x*x squares (1 ulp in ~0.1 % of squares).
CANON_TOL = 1e-11


@pytest.mark.parametrize("alg", REFERENCE_ALGORITHMS)
@pytest.mark.parametrize("sp", ["p", "e"])
@pytest.mark.parametrize("nsteps,rec", [(200, 50), (200, 30)])
def test_track_batch_uniform(golden_dir, alg, sp, nsteps, rec):
    g = load(golden_dir, "track_batch_uniform.npz")
    eF, bF = _uniform_fields(g)
    tag = f"{sp}_{alg}_{nsteps}_{rec}"
    args = (g[f"r0_{sp}"], g[f"v0_{sp}"], eF, bF, nsteps, float(g[f"dt_{sp}"]), rec)
    rh, vh, act = OraclePusher(float(g[f"qm_{sp}"]), alg, as_shipped=True).track_batch(*args)
    bitwise(rh, g[f"r_{tag}"])
    bitwise(vh, g[f"v_{tag}"])
    bitwise(act, g[f"a_{tag}"])
    rh, vh, _ = OraclePusher(float(g[f"qm_{sp}"]), alg).track_batch(*args)
    assert rel_err(rh, g[f"r_{tag}"]) < CANON_TOL
    assert rel_err(vh, g[f"v_{tag}"]) < CANON_TOL


@pytest.mark.parametrize("alg", REFERENCE_ALGORITHMS)
@pytest.mark.parametrize("sp", ["p", "e"])
def test_track_batch_map(golden_dir, alg, sp):
    g = load(golden_dir, "track_batch_map.npz")
    grid, E, B = _map_fields(g, None)
    v0 = g["v0p"] if sp == "p" else g["v0e"]
    qm, dt = float(g[f"qm_{sp}"]), float(g[f"dt_{sp}"])
    for interp, exact in (("coordmapper", True), ("numba", False)):
        eF = OField.from_arrays(grid, E, scaling=1.0, interp=interp)
        bF = OField.from_arrays(grid, B, scaling=2.0, interp=interp)
        rh, vh, _ = OraclePusher(qm, alg, as_shipped=exact).track_batch(g["r0"], v0, eF, bF, 120, dt, 40)
        if exact:
            bitwise(rh, g[f"r_{sp}_{alg}"])
            bitwise(vh, g[f"v_{sp}_{alg}"])
        else:
            assert rel_err(rh, g[f"r_{sp}_{alg}"]) < CANON_TOL
            assert rel_err(vh, g[f"v_{sp}_{alg}"]) < CANON_TOL


@pytest.mark.parametrize("alg", REFERENCE_ALGORITHMS)
def test_track_push_single(golden_dir, alg):
    g = load(golden_dir, "track_single.npz")
    gu = load(golden_dir, "track_batch_uniform.npz")
    eF, bF = _uniform_fields(gu)
    pu = OraclePusher(float(g["qm"]), alg, as_shipped=True)
    rh, vh = pu.track(g["r0"][3], g["v0"][3], eF, bF, 50, 1e-11, 7)
    bitwise(rh, g[f"trk_r_{alg}"])
    bitwise(vh, g[f"trk_v_{alg}"])
    rn, vn = pu.push(g["r0"][5], g["v0"][5], eF, bF, 1e-11)
    bitwise(rn, g[f"push_r_{alg}"])
    bitwise(vn, g[f"push_v_{alg}"])
    rn, vn = pu.push_batch(g["r0"], g["v0"], eF, bF, 1e-11)
    bitwise(rn, g[f"pb_r_{alg}"])
    bitwise(vn, g[f"pb_v_{alg}"])


class _Wall:
    def __init__(self, wall):
        self.wall = wall

    def segment_intersects_surface(self, r_old, r_new):
        return {"hit_mask": (r_old[:, 0] <= self.wall) & (r_new[:, 0] > self.wall)}


@pytest.mark.parametrize("alg", ["boris", "rk4", "vay_rel"])
def test_track_batch_particle_loss(golden_dir, alg):
    g = load(golden_dir, "track_batch_loss.npz")
    eF = OField.constant(0.0, 0.0, 0.0)
    bF = OField.constant(0.0, 0.0, 0.2)
    wall = _Wall(float(g["wall"]))
    for as_shipped in (True, False):
        pu = OraclePusher(float(g["qm"]), alg, as_shipped=as_shipped)
        rh, vh, act = pu.track_batch(g["r0"], g["v0"], eF, bF, 60, 1e-10, 10, electrode_assembly=wall)
        if as_shipped:
            bitwise(rh, g[f"r_{alg}"])
            bitwise(vh, g[f"v_{alg}"])
        else:
            assert rel_err(rh, g[f"r_{alg}"]) < CANON_TOL
            assert rel_err(vh, g[f"v_{alg}"]) < CANON_TOL
        bitwise(act, g[f"a_{alg}"])


def test_track_batch_middle_loss_quirk(golden_dir):
    """The reference's no-numba Boris finalisation raises when a non-final particle is lost."""
    g = load(golden_dir, "track_batch_loss.npz")
    assert str(g["middle_loss_error"]) == "IndexError"
    eF = OField.constant(0.0, 0.0, 0.0)
    bF = OField.constant(0.0, 0.0, 0.2)
    pu = OraclePusher(float(g["qm"]), "boris", as_shipped=True)
    with pytest.raises(IndexError):
        pu.track_batch(g["r0"], g["v0_middle"], eF, bF, 60, 1e-10, 10, electrode_assembly=_Wall(float(g["wall"])))
