"""The workloads at their full sizes and horizons against the CPU oracle.

* C2: 1000 protons spread over the 1e6-particle bunch, through the 128^3 map (periodic in z) for the
  full 1e4 steps: strict FP64 equals the canonical oracle bit for bit, fp64_fast within 1e-11 and
  fp32_mixed within 1e-6 of the path length (the precision bar of the FP32-mixed form).
* C3: 500 of the 4e6 electrons through the 96x96x256 TM010-like map with its RF factors for the full
  5000 rk4_rel steps: the same bars, and the FP32-mixed energy error.
* C4: all 512 bunches (4.1e7 particles) in the batched launch.  Every bunch of a stratified sample
  (each species x algorithm: the smallest and largest scale and step count) is checked on 256 of its
  particles against the oracle's track_batch over its full step count with its ScaledField factors
  (pusher.py:894-1012, field.py:661-690); bunches of the sample run alone equal the batch bit for bit.
* C5: a mix of C5 jobs (N log-uniform 1e3-2e6, steps 1e2-2e4, boris / vay / rk4, four maps) through the
  continuous-batching server with a pool too small for all of them at once; a sample of jobs per
  algorithm and map checked on 256 particles against the oracle; jobs run alone equal the mix.
* Particles parked in the periodic map's wrap cell (z between the last node and z0 + L).

The oracle runs in spawned worker processes (tests/oracle_pool.py) while the GPU works.
"""
import numpy as np
import pytest

import pypatools_amd as pa
from oracle import kernels as K
from oracle.field import OPeriodicZ
from oracle.pusher import OraclePusher
import oracle_pool
from oracle_pool import OraclePool, c5_map_scales

pytestmark = pytest.mark.gpu

PARTICLES = 256          # oracle subset per bunch / job


@pytest.fixture(scope="module", autouse=True)
def _device():
    from pypatools_amd import device
    if device.device_count() == 0:
        pytest.skip("no HIP device visible")


@pytest.fixture(scope="module")
def pool():
    p = OraclePool()
    yield p
    p.close()


@pytest.fixture(scope="module")
def c2_fields():
    from pypatools_amd.workloads import c2_map
    grid, E, B = c2_map(128)
    return grid, E, B, pa.Field.from_arrays(grid, E, periodic_z=True), pa.Field.from_arrays(grid, B, periodic_z=True)


def path_err(r, r_ref, v0, nsteps, dt):
    """max |dr| / (|v0| N dt) over particles (the NaN-free last records)."""
    path = np.linalg.norm(v0, axis=1) * nsteps * dt
    return float(np.max(np.linalg.norm(r - r_ref, axis=1) / path))


def gamma(v):
    return 1.0 / np.sqrt(1.0 - np.sum(v * v, axis=1) / K.C2)


def _flush(msg):
    print(msg, flush=True)


_ORACLE = {}      # (workload, index) -> future of the oracle's last records, shared by both precisions


def _oracle(pool, key, job):
    if key not in _ORACLE:
        _ORACLE[key] = pool.ex.submit(oracle_pool.track_last, job)
    return _ORACLE[key]


# --------------------------------------------------------------------------- C2 / C3 full horizon


def test_c2_full_horizon_against_oracle(pool, c2_fields):
    from pypatools_amd.workloads import c2_bunch
    grid, E, B, fe, fb = c2_fields
    r_all, v_all = c2_bunch(1_000_000)
    r0, v0 = r_all[::1000].copy(), v_all[::1000].copy()
    N, dt = 10_000, 5e-11
    _flush("C2 full horizon: oracle on 1000 particles x 1e4 steps")
    jobs = [dict(map=("c2", 128, None), alg="boris", qm=pa.IonSpecies("proton").q_over_m, dt=dt, nsteps=N,
                 r0=r0[i], v0=v0[i]) for i in np.array_split(np.arange(len(r0)), 8)]
    futs = [pool.ex.submit(oracle_pool.track_last, j) for j in jobs]
    out = {}
    for prec in ("fp64", "fp64_fast", "fp32_mixed"):
        rh, vh, _ = pa.Pusher(pa.IonSpecies("proton"), "boris", precision=prec).track_batch(r0, v0, fe, fb, N, dt, N)
        out[prec] = (rh[-1], vh[-1])
    res = [f.result() for f in futs]
    ro, vo = np.concatenate([r for r, _ in res]), np.concatenate([v for _, v in res])
    np.testing.assert_array_equal(out["fp64"][0], ro)
    np.testing.assert_array_equal(out["fp64"][1], vo)
    e_fast = path_err(out["fp64_fast"][0], ro, v0, N, dt)
    e_mixed = path_err(out["fp32_mixed"][0], ro, v0, N, dt)
    _flush(f"C2 1e4 steps: fp64_fast {e_fast:.2e}, fp32_mixed {e_mixed:.2e} of the path")
    assert e_fast < 1e-11, e_fast
    assert e_mixed < 1e-6, e_mixed
    half = 0.0635
    assert np.all((np.abs(ro[:, 0]) <= half) & (np.abs(ro[:, 1]) <= half))      # the whole track inside the map


def test_c3_full_horizon_against_oracle(pool):
    from pypatools_amd.workloads import c3_bunch, c3_map
    grid, E, B, w = c3_map()
    r_all, v_all, phase = c3_bunch(4_000_000)
    r0, v0 = r_all[::8000].copy(), v_all[::8000].copy()
    N, dt = 5000, 2e-13
    qm = pa.IonSpecies("electron").q_over_m
    _flush("C3 full horizon: oracle on 500 electrons x 5000 rk4_rel steps")
    futs = [pool.ex.submit(oracle_pool.track_last, dict(map=("c3",), alg="rk4_rel", qm=qm, dt=dt, nsteps=N, r0=r0[i],
                                                v0=v0[i], rf=(w, phase)))
            for i in np.array_split(np.arange(len(r0)), 8)]
    fe = pa.RFField(pa.Field.from_arrays(grid, E), w, phase, "cos")
    fb = pa.RFField(pa.Field.from_arrays(grid, B), w, phase, "sin")
    out = {}
    for prec in ("fp64", "fp64_fast", "fp32_mixed"):
        rh, vh, _ = pa.Pusher(pa.IonSpecies("electron"), "rk4_rel", precision=prec).track_batch(r0, v0, fe, fb, N,
                                                                                              dt, N)
        out[prec] = (rh[-1], vh[-1])
    res = [f.result() for f in futs]
    ro, vo = np.concatenate([r for r, _ in res]), np.concatenate([v for _, v in res])
    np.testing.assert_array_equal(out["fp64"][0], ro)
    np.testing.assert_array_equal(out["fp64"][1], vo)
    e_fast = path_err(out["fp64_fast"][0], ro, v0, N, dt)
    e_mixed = path_err(out["fp32_mixed"][0], ro, v0, N, dt)
    go = gamma(vo)
    e_energy = float(np.max(np.abs(gamma(out["fp32_mixed"][1]) - go) / (go - 1.0)))
    gain_kev = float(np.mean(go - gamma(v0)) * 510.99895069)
    z_end = ro[:, 2]
    _flush(f"C3 5000 steps: fp64_fast {e_fast:.2e}, fp32_mixed {e_mixed:.2e} of the path, energy {e_energy:.2e}; "
           f"mean gain {gain_kev:.0f} keV, z {z_end.min():.4f}..{z_end.max():.4f} m")
    assert e_fast < 1e-11, e_fast
    assert e_mixed < 1e-6, e_mixed
    assert e_energy < 1e-5, e_energy
    assert np.all(z_end < grid["z"][-1]) and np.all(np.hypot(ro[:, 0], ro[:, 1]) < grid["x"][-1])   # inside the map
    assert gain_kev > 1000.0                                                  # the RF field accelerates them


# --------------------------------------------------------------------------- C4 batched, full size


def _c4_sample(bunches):
    """Indices of the smallest / largest scale and step count of each (species, algorithm) group."""
    groups = {}
    for k, b in enumerate(bunches):
        groups.setdefault((b.q_over_m, b.algorithm), []).append(k)
    pick = set()
    for ks in groups.values():
        for key in (lambda k: bunches[k].escale, lambda k: bunches[k].nsteps):
            pick.add(min(ks, key=key))
            pick.add(max(ks, key=key))
    return sorted(pick), len(groups)


@pytest.fixture(scope="module")
def c4_bunch_list():
    from pypatools_amd.batch import c4_bunches
    return c4_bunches()


@pytest.mark.parametrize("prec,tol", [("fp32_mixed", 1e-6), ("fp64_fast", 1e-11)])
def test_c4_all_bunches_against_oracle_and_alone(pool, c2_fields, c4_bunch_list, prec, tol):
    from pypatools_amd.batch import BatchedTracker
    grid, E, B, fe, fb = c2_fields
    bunches = c4_bunch_list
    assert len(bunches) == 512 and sum(b.n for b in bunches) > 3e7
    sample, ngroups = _c4_sample(bunches)
    assert ngroups == 6
    _flush(f"C4 {prec}: 512 bunches on the GPU, oracle on {len(sample)} sampled bunches x {PARTICLES}")
    futs = [_oracle(pool, ("c4", k), dict(map=("c2", 128, None), alg=bunches[k].algorithm, qm=bunches[k].q_over_m,
                                          dt=bunches[k].dt, nsteps=bunches[k].nsteps, r0=bunches[k].r0[:PARTICLES],
                                          v0=bunches[k].v0[:PARTICLES], escale=bunches[k].escale,
                                          bscale=bunches[k].bscale))
            for k in sample]
    bt = BatchedTracker([(fe, fb)], prec).load(bunches)
    bt.run()
    res = bt.results()
    worst = 0.0
    for k, f in zip(sample, futs):
        ro, vo = f.result()
        b = bunches[k]
        e = path_err(res[k][0][:PARTICLES], ro, b.v0[:PARTICLES], b.nsteps, b.dt)
        ev = float(np.max(np.linalg.norm(res[k][1][:PARTICLES] - vo, axis=1) / np.linalg.norm(vo, axis=1)))
        worst = max(worst, e)
        assert e < tol, (k, b.algorithm, b.nsteps, b.escale, e)
        assert ev < 10 * tol, (k, b.algorithm, ev)
    _flush(f"C4 {prec}: worst path error {worst:.2e} over {len(sample)} bunches")
    for k in sorted(sample, key=lambda k: bunches[k].n)[:3]:       # alone = in the mix, bit for bit
        one = BatchedTracker([(fe, fb)], prec).load([bunches[k]])
        one.run()
        (r1, v1), = one.results()
        np.testing.assert_array_equal(res[k][0], r1)
        np.testing.assert_array_equal(res[k][1], v1)


# --------------------------------------------------------------------------- C5 served, full size


@pytest.fixture(scope="module")
def c5_fields():
    from pypatools_amd.workloads import c2_map
    grid, E, B = c2_map(128)
    out = []
    for k in range(4):
        se, sb = c5_map_scales(k)
        out.append((pa.Field.from_arrays(grid, {c: se * E[c] for c in "xyz"}, periodic_z=True),
                    pa.Field.from_arrays(grid, {c: sb * B[c] for c in "xyz"}, periodic_z=True)))
    return out


def _c5_sample(jobs, rk4_max_steps=6000):
    """Per algorithm, up to 4 jobs on distinct maps (RK4 jobs of at most rk4_max_steps, for the oracle's time)."""
    pick = []
    for alg in ("boris", "vay", "rk4"):
        seen = set()
        for k, j in enumerate(jobs):
            if j.algorithm != alg or j.map_index in seen or (alg == "rk4" and j.nsteps > rk4_max_steps):
                continue
            seen.add(j.map_index)
            pick.append(k)
    return pick


@pytest.mark.parametrize("prec,tol", [("fp32_mixed", 1e-6), ("fp64_fast", 1e-11)])
def test_c5_served_mix_against_oracle_and_alone(pool, c5_fields, prec, tol):
    from pypatools_amd.serving import TrackingServer, c5_jobs
    jobs = c5_jobs(48, 1e9, seed=21)
    sample = _c5_sample(jobs)
    assert {jobs[k].algorithm for k in sample} == {"boris", "vay", "rk4"} and len(sample) >= 9
    _flush(f"C5 {prec}: {len(jobs)} jobs through the server, oracle on {len(sample)} sampled jobs x {PARTICLES}")
    futs = [_oracle(pool, ("c5", k), dict(map=("c2", 128, jobs[k].map_index), alg=jobs[k].algorithm,
                                          qm=jobs[k].q_over_m, dt=jobs[k].dt, nsteps=jobs[k].nsteps,
                                          r0=jobs[k].r0[:PARTICLES], v0=jobs[k].v0[:PARTICLES]))
            for k in sample]
    total = sum(j.n for j in jobs)
    srv = TrackingServer(c5_fields, prec, pool_particles=max(total // 3, max(j.n for j in jobs)))
    for j in jobs:
        srv.submit(j)
    srv.run_until_idle()
    assert all(j.state == "done" for j in jobs)
    assert srv.particle_steps == sum(j.n * j.nsteps for j in jobs)
    worst = 0.0
    for k, f in zip(sample, futs):
        ro, vo = f.result()
        j = jobs[k]
        e = path_err(j.result[0][:PARTICLES], ro, j.v0[:PARTICLES], j.nsteps, j.dt)
        worst = max(worst, e)
        assert e < tol, (k, j.algorithm, j.nsteps, j.map_index, e)
    _flush(f"C5 {prec}: worst path error {worst:.2e} over {len(sample)} jobs; {srv.rounds} rounds")
    for k in sorted(sample, key=lambda k: jobs[k].n)[:3]:
        j = jobs[k]
        alone = type(j)(j.r0, j.v0, j.q_over_m, j.algorithm, j.nsteps, j.dt, j.map_index)
        one = TrackingServer(c5_fields, prec, pool_particles=j.n)
        one.submit(alone)
        one.run_until_idle()
        np.testing.assert_array_equal(alone.result[0], j.result[0])
        np.testing.assert_array_equal(alone.result[1], j.result[1])


# --------------------------------------------------------------------------- periodic wrap cell


def test_periodic_wrap_cell_on_device(c2_fields):
    """Field values and a slow track inside the wrap cell (z between z_last and z0 + L, and below z0)."""
    grid, E, B, fe, fb = c2_fields
    rng = np.random.default_rng(41)
    M = 4000
    dz = grid["z"][1] - grid["z"][0]
    L = len(grid["z"]) * dz
    z_last = grid["z"][-1]
    f = rng.uniform(0.0, 1.0, M)
    pts = np.column_stack([rng.normal(0, 5e-3, M), rng.normal(0, 5e-3, M),
                           np.where(np.arange(M) % 2 == 0, z_last + f * dz, grid["z"][0] - f * dz + 3 * L)])
    oe, ob = OPeriodicZ(grid, E), OPeriodicZ(grid, B)
    np.testing.assert_array_equal(fe(pts), oe(pts))
    np.testing.assert_array_equal(fb(pts), ob(pts))
    # 1 keV protons drift 0.022 mm per step from 0-0.5 mm past z_last: most of the first 25 steps are
    # taken inside the wrap cell, and the rest of the 100 in the map's first cells after the wrap
    speed = np.sqrt(2 * 1e-3 / 938.272) * K.CLIGHT
    M2 = 512
    r0 = np.column_stack([rng.normal(0, 3e-3, M2), rng.normal(0, 3e-3, M2), z_last + rng.uniform(0, 0.5, M2) * dz])
    v0 = np.column_stack([rng.normal(0, 0.1, M2), rng.normal(0, 0.1, M2), np.ones(M2)]) * speed
    qm = pa.IonSpecies("proton").q_over_m
    for alg in ("boris", "rk4", "vay"):
        orh, ovh, _ = OraclePusher(qm, alg).track_batch(r0, v0, oe, ob, 100, 5e-11, 25)
        rh, vh, _ = pa.Pusher(pa.IonSpecies("proton"), alg, precision="fp64").track_batch(r0, v0, fe, fb, 100, 5e-11,
                                                                                          25)
        np.testing.assert_array_equal(rh, orh)
        np.testing.assert_array_equal(vh, ovh)
        rm, _, _ = pa.Pusher(pa.IonSpecies("proton"), alg, precision="fp32_mixed").track_batch(r0, v0, fe, fb, 100,
                                                                                               5e-11, 100)
        assert path_err(rm[-1], orh[-1], v0, 100, 5e-11) < 1e-6
    z_25 = orh[1][:, 2]                       # after 25 steps
    assert np.mean((z_25 > z_last) & (z_25 < grid["z"][0] + L)) > 0.5
    assert np.all(orh[-1][:, 2] > grid["z"][0] + L)      # and all of them wrapped by step 100
