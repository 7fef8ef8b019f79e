"""Electrode loss on the device (pp_collide_segments, TriangleElectrodes) against the brute-force oracle."""
import numpy as np
import pytest

import pypatools_amd as pa
from oracle.electrodes import OracleElectrodes, segment_hits
from oracle.field import OField
from oracle.pusher import OraclePusher
from pypatools_amd.electrodes import TriangleElectrodes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _device():
    from pypatools_amd import device
    if device.device_count() == 0:
        pytest.skip("no HIP device visible")


def _soup(rng, nt=3000):
    v = rng.uniform(-1, 1, (3 * nt, 3))
    v[1::3] = v[0::3] + 0.15 * rng.standard_normal((nt, 3))
    v[2::3] = v[0::3] + 0.15 * rng.standard_normal((nt, 3))
    return v, np.arange(3 * nt).reshape(nt, 3)


@pytest.mark.parametrize("kind", ["box", "cylinder", "soup", "empty", "one"])
def test_segments_match_brute_force(kind):
    rng = np.random.default_rng(50)
    if kind == "box":
        m = TriangleElectrodes.box([-0.5, -0.4, -0.3], [0.6, 0.5, 0.7])
    elif kind == "cylinder":
        m = TriangleElectrodes.cylinder(0.6, -1.0, 1.0, n_phi=48, n_z=12)
    elif kind == "soup":
        m = TriangleElectrodes(*_soup(rng))
    elif kind == "empty":
        m = TriangleElectrodes(np.zeros((0, 3)), np.zeros((0, 3), dtype=np.int32))
    else:
        m = TriangleElectrodes([[0, 0, 0], [1, 0, 0], [0, 1, 0]], [[0, 1, 2]])
    n = 5000
    r0 = rng.uniform(-1.2, 1.2, (n, 3))
    r1 = r0 + 0.3 * rng.standard_normal((n, 3))
    r1[:50] = r0[:50]                                   # zero-length segments
    got = m.segment_intersects_surface(r0, r1)["hit_mask"]
    want = segment_hits(m.vertices, m.triangles, r0, r1)
    np.testing.assert_array_equal(got, want)
    if kind in ("box", "cylinder", "soup"):
        assert 0.05 < want.mean() < 0.95
    assert not np.any(got[:50] & ~want[:50])


def test_listed_and_crowded_cells_match_brute_force():
    """Short segments are decided by the level-0 triangle lists where every cell they cover has one, and by
    the hierarchy walk elsewhere.  Here 160 large triangles crowd the left half (about a quarter of its cells
    are overlapped by more than the 32 triangles a list holds, so segments there cover listed cells, cells
    without a list, or both), small triangles fill the right half, and short random segments land on either
    side and across the middle: the device's hit masks equal the brute force."""
    rng = np.random.default_rng(63)
    nb = 160
    big = np.column_stack([rng.uniform(-1.0, -0.1, (3 * nb, 1)), rng.uniform(-1.0, 1.0, (3 * nb, 2))])
    vs, ts = _soup(rng, 2000)
    vs = vs * np.array([0.45, 1.0, 1.0]) + np.array([0.55, 0.0, 0.0])
    v = np.vstack([big, vs])
    t = np.vstack([np.arange(3 * nb).reshape(nb, 3), ts + 3 * nb]).astype(np.int32)
    m = TriangleElectrodes(v, t)
    n = 20000
    r0 = rng.uniform(-1.1, 1.1, (n, 3))
    r1 = r0 + 0.05 * rng.standard_normal((n, 3))
    got = m.segment_intersects_surface(r0, r1)["hit_mask"]
    want = segment_hits(m.vertices, m.triangles, r0, r1)
    np.testing.assert_array_equal(got, want)
    left, right = r0[:, 0] < -0.2, r0[:, 0] > 0.2
    assert want[left].mean() > 0.05 and want[right].mean() > 0.05


def test_aimed_segments_over_many_candidate_cells_match_brute_force():
    """A 12^3 lattice of small triangles (circumradius 0.01, normal (1,1,1)) keeps each candidate cell's
    subtree narrow, so a cell left out of a segment's candidate set loses the hit.  Segments are aimed at
    a triangle (or just beside it) along axis, face-diagonal and body-diagonal directions with lengths up to
    0.2: their boxes cover 1..8 cells in every shape (1x1xk, 1xkx1, kx1x1, 2x2x1, 2x2x2, ...) at the first
    levels of the candidate pyramid."""
    rng = np.random.default_rng(61)
    g = (np.arange(12) + 0.5) / 12.0
    cen = np.stack(np.meshgrid(g, g, g, indexing="ij"), axis=-1).reshape(-1, 3)
    e1 = np.array([1.0, -1.0, 0.0]) / np.sqrt(2.0)
    e2 = np.array([1.0, 1.0, -2.0]) / np.sqrt(6.0)
    ang = np.array([0.0, 2.0 * np.pi / 3.0, 4.0 * np.pi / 3.0])
    corners = 0.01 * (np.cos(ang)[:, None] * e1 + np.sin(ang)[:, None] * e2)       # (3, 3)
    v = (cen[:, None, :] + corners[None]).reshape(-1, 3)
    m = TriangleElectrodes(v, np.arange(len(v)).reshape(-1, 3))
    n = 8000
    dirs = np.array([[1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 0], [1, 0, 1], [0, 1, 1], [1, 1, 1]], float)
    d = dirs[rng.integers(0, len(dirs), n)] + 0.05 * rng.standard_normal((n, 3))
    d *= rng.choice([-1.0, 1.0], (n, 1))
    d /= np.linalg.norm(d, axis=1)[:, None]
    t = rng.integers(0, len(cen), n)
    off = rng.uniform(0.0, 0.02, n)[:, None] * np.where(rng.random((n, 1)) < 0.5, e1, e2)   # beside it
    p = cen[t] + off
    r0 = p - rng.uniform(0.001, 0.1, (n, 1)) * d
    r1 = p + rng.uniform(0.001, 0.1, (n, 1)) * d
    got = m.segment_intersects_surface(r0, r1)["hit_mask"]
    want = segment_hits(m.vertices, m.triangles, r0, r1)
    np.testing.assert_array_equal(got, want)
    assert 0.2 < want.mean() < 0.9


@pytest.mark.parametrize("kind", ["box", "cylinder", "soup"])
def test_short_segments_near_surfaces_match_brute_force(kind):
    """Segments of 1e-7 .. 1e-2 of the mesh size around the surface (the occupancy grid decides most of
    them; endpoints on vertices, on edges and exactly on faces included) agree with the brute force."""
    rng = np.random.default_rng(53)
    if kind == "box":
        m = TriangleElectrodes.box([-0.5, -0.4, -0.3], [0.6, 0.5, 0.7])
    elif kind == "cylinder":
        m = TriangleElectrodes.cylinder(0.6, -1.0, 1.0, n_phi=48, n_z=12)
    else:
        m = TriangleElectrodes(*_soup(rng))
    v, t = m.vertices, m.triangles
    n = 20000
    k = rng.integers(0, len(t), n)
    bary = rng.dirichlet([1, 1, 1], n)
    bary[: n // 10] = np.eye(3)[rng.integers(0, 3, n // 10)]             # on a vertex
    e = slice(n // 10, n // 5)
    bary[e, 0] = 0.0
    bary[e] /= bary[e].sum(axis=1, keepdims=True)                       # on an edge
    on = np.einsum("nk,nkd->nd", bary, v[t[k]])
    size = float(np.max(v.max(axis=0) - v.min(axis=0)))
    length = size * 10.0 ** rng.uniform(-7, -2, n)
    d = rng.standard_normal((n, 3))
    d *= (length / np.linalg.norm(d, axis=1))[:, None]
    frac = rng.uniform(-1.5, 0.5, n)                                    # where the surface point sits on the segment
    r0 = on + frac[:, None] * d
    r1 = r0 + d
    r1[: n // 20] = on[: n // 20]                                       # ending exactly on the surface point
    got = m.segment_intersects_surface(r0, r1)["hit_mask"]
    want = segment_hits(v, t, r0, r1)
    np.testing.assert_array_equal(got, want)
    assert 0.05 < want.mean() < 0.95


def test_mesh_validation():
    with pytest.raises(ValueError):
        TriangleElectrodes([[0, 0, 0], [1, 0, 0]], [[0, 1, 2]])
    with pytest.raises(ValueError):
        TriangleElectrodes([[0, 0, np.nan], [1, 0, 0], [0, 1, 0]], [[0, 1, 2]])
    info = TriangleElectrodes.cylinder(1.0, 0.0, 1.0, 32, 8).info()
    assert info["triangles"] == 512 and info["nodes"] >= 512 // 4


@pytest.mark.parametrize("alg", ["boris", "rk4", "vay_rel"])
def test_track_with_device_losses_equals_host_callback_and_oracle(alg):
    """Protons gyrating in 0.5 T inside a 6 mm beam pipe: the large orbits hit the pipe.  The device loss
    loop (TriangleElectrodes), the host-callback loop (the same mesh through the reference's protocol)
    and the oracle pusher agree bit for bit, histories and active masks."""
    rng = np.random.default_rng(51)
    n = 600
    pipe = TriangleElectrodes.cylinder(6e-3, -0.05, 0.05, n_phi=40, n_z=10)
    orc_e = OracleElectrodes(pipe.vertices, pipe.triangles)
    r0 = np.column_stack([rng.uniform(-2e-3, 2e-3, n), rng.uniform(-2e-3, 2e-3, n), np.zeros(n)])
    v0 = np.column_stack([rng.normal(0, 3e5, n), rng.normal(0, 3e5, n), rng.normal(0, 1e5, n)])
    E, B = pa.Field(dim=0, field={"x": 0.0, "y": 0.0, "z": 0.0}), pa.Field(dim=0, field={"x": 0.0, "y": 0.0, "z": 0.5})
    qm = pa.IonSpecies("proton").q_over_m
    pusher = pa.Pusher(pa.IonSpecies("proton"), alg, electrode_assembly=pipe)
    rh, vh, act = pusher.track_batch(r0, v0, E, B, 400, 2e-10, 50)
    p2 = pa.Pusher(pa.IonSpecies("proton"), alg, electrode_assembly=orc_e)
    rh2, vh2, act2 = p2.track_batch(r0, v0, E, B, 400, 2e-10, 50)
    np.testing.assert_array_equal(act, act2)
    np.testing.assert_array_equal(rh, rh2)
    np.testing.assert_array_equal(vh, vh2)
    orc = OraclePusher(qm, alg)
    orh, ovh, oact = orc.track_batch(r0, v0, OField.constant(0, 0, 0), OField.constant(0, 0, 0.5), 400, 2e-10, 50,
                                     electrode_assembly=orc_e)
    np.testing.assert_array_equal(act, oact)
    np.testing.assert_array_equal(rh, orh)
    np.testing.assert_array_equal(vh, ovh)
    assert 0.05 < 1 - act.mean() < 0.99                  # some, not all, particles were lost


@pytest.mark.parametrize("prec", ["fp64", "fp64_fast", "fp32_mixed"])
@pytest.mark.parametrize("alg", ["boris", "rk4_rel", "vay"])
def test_lossy_track_through_a_map_with_compaction(alg, prec):
    """A beam pipe inside a 3-D map: the fused loss test with compaction every 7 steps (30 compactions,
    each a partition plus cell sort) equals the host-callback route step by step, histories and active
    masks bit for bit (in FP32-mixed and FP64-fast both routes run the hot kernel's loss build); in
    strict FP64 it also equals the oracle pusher with the oracle mesh, and in the fast precisions it
    follows the oracle to their tolerance: the same particles are lost (a particle whose segment grazes
    the wall may go either way: at most 0.5 % differ) and the survivors' final positions agree to 1e-6
    (FP32-mixed) or 1e-11 (FP64-fast) of the path."""
    rng = np.random.default_rng(52)
    n = 2500
    x = (np.arange(16) - 7.5) * 1e-3
    z = np.arange(16) * 1e-3 - 2e-3
    grid = {"x": x, "y": x, "z": z}
    X, Y, Z = np.meshgrid(x, x, z, indexing="ij")
    E = {"x": 3e5 * X / 7.5e-3, "y": 3e5 * Y / 7.5e-3, "z": 1e5 + 0 * Z}
    B = {"x": 0.02 * Y / 7.5e-3, "y": 0.02 * X / 7.5e-3, "z": 0.4 + 0.1 * np.cos(Z / 2e-3)}
    fe, fb = pa.Field.from_arrays(grid, E), pa.Field.from_arrays(grid, B)
    pipe = TriangleElectrodes.cylinder(5e-3, -2e-3, 13e-3, n_phi=36, n_z=15)
    orc_e = OracleElectrodes(pipe.vertices, pipe.triangles)
    r0 = np.column_stack([rng.uniform(-3e-3, 3e-3, n), rng.uniform(-3e-3, 3e-3, n), rng.uniform(0, 2e-3, n)])
    v0 = np.column_stack([rng.normal(0, 4e5, n), rng.normal(0, 4e5, n), rng.uniform(1e5, 6e5, n)])
    ion = pa.IonSpecies("proton")
    dev = pa.Pusher(ion, alg, precision=prec, electrode_assembly=pipe, sort_every=7)
    rh, vh, act = dev.track_batch(r0, v0, fe, fb, 210, 5e-11, 30)
    host = pa.Pusher(ion, alg, precision=prec, electrode_assembly=orc_e)
    rh2, vh2, act2 = host.track_batch(r0, v0, fe, fb, 210, 5e-11, 30)
    np.testing.assert_array_equal(act, act2)
    np.testing.assert_array_equal(rh, rh2)
    np.testing.assert_array_equal(vh, vh2)
    assert 0.1 < 1 - act.mean() < 0.95
    orc = OraclePusher(ion.q_over_m, alg)
    # particles are independent: the oracle runs on the first k of them (all for the bit-for-bit FP64 check)
    k = n if prec == "fp64" else 800
    orh, ovh, oact = orc.track_batch(r0[:k], v0[:k], OField.from_arrays(grid, E), OField.from_arrays(grid, B), 210,
                                     5e-11, 30, electrode_assembly=orc_e)
    if prec == "fp64":
        np.testing.assert_array_equal(act, oact)
        np.testing.assert_array_equal(rh, orh)
        np.testing.assert_array_equal(vh, ovh)
    else:
        assert np.mean(act[:k] != oact) <= 5e-3, np.mean(act[:k] != oact)
        both = act[:k] & oact
        assert 0.1 < 1 - oact.mean() < 0.95
        path = np.linalg.norm(v0[:k][both], axis=1) * 210 * 5e-11
        err = np.max(np.linalg.norm(rh[-1][:k][both] - orh[-1][both], axis=1) / path)
        assert err < (1e-6 if prec == "fp32_mixed" else 1e-11), err


def test_every_particle_lost_ends_the_run_early():
    """All particles leave a small closed box within a few steps: the device path stops launching when no
    slot is live; histories after each loss stay NaN and equal the host-callback route and the oracle."""
    rng = np.random.default_rng(54)
    n = 3000
    box = TriangleElectrodes.box([-1e-3, -1e-3, -1e-3], [1e-3, 1e-3, 1e-3])
    orc_e = OracleElectrodes(box.vertices, box.triangles)
    r0 = rng.uniform(-5e-4, 5e-4, (n, 3))
    v0 = rng.standard_normal((n, 3)) * 2e6
    E, B = pa.Field(dim=0, field={"x": 0.0, "y": 0.0, "z": 0.0}), pa.Field(dim=0, field={"x": 0.0, "y": 0.0, "z": 0.1})
    ion = pa.IonSpecies("proton")
    rh, vh, act = pa.Pusher(ion, "boris", electrode_assembly=box, sort_every=3).track_batch(r0, v0, E, B, 200, 1e-10, 5)
    rh2, vh2, act2 = pa.Pusher(ion, "boris", electrode_assembly=orc_e).track_batch(r0, v0, E, B, 200, 1e-10, 5)
    assert not act.any()
    np.testing.assert_array_equal(act, act2)
    np.testing.assert_array_equal(rh, rh2)
    np.testing.assert_array_equal(vh, vh2)
    orh, ovh, oact = OraclePusher(ion.q_over_m, "boris").track_batch(
        r0, v0, OField.constant(0, 0, 0), OField.constant(0, 0, 0.1), 200, 1e-10, 5, electrode_assembly=orc_e)
    np.testing.assert_array_equal(rh, orh)
    assert np.all(np.isnan(rh[-1]))


def test_merged_electrodes_collide_like_their_parts():
    """A merged assembly (box + tube) hits exactly where either part does."""
    rng = np.random.default_rng(55)
    a = TriangleElectrodes.box([-0.3, -0.3, 0.2], [0.3, 0.3, 0.5])
    b = TriangleElectrodes.cylinder(0.8, -1.0, 1.0, n_phi=32, n_z=8)
    m = TriangleElectrodes.merge(a, b)
    r0 = rng.uniform(-1, 1, (4000, 3))
    r1 = r0 + 0.4 * rng.standard_normal((4000, 3))
    hm = m.segment_intersects_surface(r0, r1)["hit_mask"]
    ha = a.segment_intersects_surface(r0, r1)["hit_mask"]
    hb = b.segment_intersects_surface(r0, r1)["hit_mask"]
    np.testing.assert_array_equal(hm, ha | hb)
    assert ha.sum() > 0 and hb.sum() > 0
