"""Batched persistent launch, HIP-graph replay and the continuous-batching server.

Particles are independent, so within one kernel every bunch / job must come
out bit for bit the same whether it runs alone or batched with others, eagerly
or from a replayed graph, admitted early or late into the server.  Against the
single-bunch hot kernel (a different kernel, different FMA contraction) the
batched results agree to the precision's tolerance.
"""
import numpy as np
import pytest

import pypatools_amd as pa
from pypatools_amd.batch import BatchedTracker, Bunch
from pypatools_amd.engine import FieldPair, ParticleState, Tracker
from pypatools_amd.serving import Job, TrackingServer
from pypatools_amd.workloads import c2_map

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _device():
    from pypatools_amd import device
    if device.device_count() == 0:
        pytest.skip("no HIP device visible")


def bitwise(a, b):
    np.testing.assert_array_equal(np.asarray(a), np.asarray(b))


@pytest.fixture(scope="module")
def maps():
    grid, E, B = c2_map(32)
    grid2 = grid
    E2 = {c: 0.5 * E[c] for c in "xyz"}
    B2 = {c: B[c][:, :, ::-1].copy() for c in "xyz"}
    f = [(pa.Field.from_arrays(grid, E, periodic_z=True), pa.Field.from_arrays(grid, B, periodic_z=True)),
         (pa.Field.from_arrays(grid2, E2, periodic_z=True), pa.Field.from_arrays(grid2, B2, periodic_z=True))]
    return grid, (E, B), (E2, B2), f


def _bunches(rng, n_list, algs):
    out = []
    sp = [("proton", 2.19e7, 5e-11), ("electron", 1.64e8, 6e-12), ("H2_1+", 9.8e6, 1e-10)]
    for k, (n, alg) in enumerate(zip(n_list, algs)):
        name, v, dt = sp[k % 3]
        r0 = np.column_stack([rng.normal(0, 2e-3, n), rng.normal(0, 2e-3, n), rng.normal(16e-3, 2e-3, n)])
        v0 = np.column_stack([rng.normal(0, 1e3, n), rng.normal(0, 1e3, n), v * (1 + 1e-3 * rng.standard_normal(n))])
        out.append(Bunch(r0, v0, pa.IonSpecies(name).q_over_m, alg, 150 + 37 * k, dt, 0.5 + 0.2 * k,
                         1.4 - 0.15 * k, k % 2))
    return out


def _alone(b, grid, fields, prec):
    """The bunch through Tracker.track (track_batch semantics, one record at the end)."""
    E, B = fields
    fe = pa.Field.from_arrays(grid, E, periodic_z=True, scaling=b.escale)
    fb = pa.Field.from_arrays(grid, B, periodic_z=True, scaling=b.bscale)
    rh, vh, _ = Tracker(b.algorithm, prec, b.q_over_m).track(b.r0, b.v0, fe, fb, b.nsteps, b.dt, b.nsteps)
    return rh[-1], vh[-1]


@pytest.mark.parametrize("prec,tol", [("fp32_mixed", 1e-6), ("fp64_fast", 1e-12)])
def test_batched_matches_each_bunch_alone(maps, prec, tol):
    """Batched launch vs the single-bunch hot kernel.  The two are different kernels of the
    fast translation unit, where the compiler chooses FMA contraction per kernel, so they agree
    to rounding (the stated tolerance of the precision), not bit for bit; bit-for-bit equality
    alone vs in a batch holds within one kernel (test_server_job_alone_equals_in_the_mix)."""
    grid, f1, f2, fields = maps
    rng = np.random.default_rng(31)
    bunches = _bunches(rng, [5000, 300, 7000, 64, 4100, 1000, 257],
                       ["boris", "vay", "rk4", "boris_rel", "vay_rel", "rk4_rel", "leapfrog"])
    bt = BatchedTracker(fields, prec, steps_per_launch=100, sort_every=120).load(bunches)
    bt.run()
    res = bt.results()
    for b, (r, v) in zip(bunches, res):
        ra, va = _alone(b, grid, f1 if b.map_index == 0 else f2, prec)
        path = np.linalg.norm(b.v0, axis=1) * b.nsteps * b.dt
        assert np.max(np.linalg.norm(r - ra, axis=1) / path) < tol, (b.algorithm, prec)
        assert np.max(np.linalg.norm(v - va, axis=1) / np.linalg.norm(va, axis=1)) < 10 * tol, (b.algorithm, prec)


def test_batched_single_bunch_equals_batched_mix(maps):
    """Within the batched kernel a bunch's result does not depend on its neighbours, bit for bit."""
    grid, f1, f2, fields = maps
    rng = np.random.default_rng(35)
    bunches = _bunches(rng, [5000, 300, 7000, 64], ["boris", "vay", "rk4", "boris_rel"])
    bt = BatchedTracker(fields, "fp32_mixed", steps_per_launch=100, sort_every=120).load(bunches)
    bt.run()
    res = bt.results()
    for k, b in enumerate(bunches):
        one = BatchedTracker(fields, "fp32_mixed", steps_per_launch=77, sort_every=50).load([b])
        one.run()
        (r1, v1), = one.results()
        bitwise(res[k][0], r1)
        bitwise(res[k][1], v1)


def test_graph_replay_equals_eager(maps):
    grid, f1, f2, fields = maps
    rng = np.random.default_rng(32)
    bunches = _bunches(rng, [6000, 500, 4500], ["boris", "vay", "rk4"])
    eager = BatchedTracker(fields, "fp32_mixed", steps_per_launch=50, sort_every=100).load(bunches)
    eager.run()
    ref = eager.results()
    g = BatchedTracker(fields, "fp32_mixed", steps_per_launch=50, sort_every=100, use_graph=True).load(bunches)
    probe = pa.Field.from_arrays(grid, f1[1], periodic_z=True)
    for chunk in (60, 90, 10 ** 6):               # several replays with eager device work in between
        g.run(nsteps=chunk)
        probe(np.zeros((1000, 3)))
    assert g.graph_count >= 1
    for (r, v), (rr, vv) in zip(g.results(), ref):
        bitwise(r, rr)
        bitwise(v, vv)


def test_server_job_alone_equals_in_the_mix(maps):
    grid, f1, f2, fields = maps
    rng = np.random.default_rng(33)
    qm = pa.IonSpecies("proton").q_over_m

    def jobs():
        r = np.random.default_rng(34)
        out = []
        for k, (n, steps, alg) in enumerate([(9000, 700, "boris"), (300, 1300, "vay"), (5000, 256, "rk4"),
                                             (1200, 900, "boris"), (20000, 300, "vay"), (64, 2000, "rk4")]):
            r0 = np.column_stack([r.normal(0, 2e-3, n), r.normal(0, 2e-3, n), r.normal(16e-3, 2e-3, n)])
            v0 = np.column_stack([np.zeros(n), np.zeros(n), 2.19e7 * (1 + 1e-3 * r.standard_normal(n))])
            out.append(Job(r0, v0, qm, alg, steps, 5e-11, k % 2))
        return out

    alone = []
    for j in jobs():
        srv = TrackingServer(fields, pool_particles=40000)
        srv.submit(j)
        srv.run_until_idle()
        alone.append(j.result)
    srv = TrackingServer(fields, pool_particles=30000)      # too small for all at once: admission waits
    mixed = jobs()
    for j in mixed[:3]:
        srv.submit(j)
    srv.round()
    srv.round()
    for j in mixed[3:]:
        srv.submit(j)
    srv.run_until_idle()
    assert all(j.state == "done" for j in mixed)
    assert srv.particle_steps == sum(j.n * j.nsteps for j in mixed)
    for j, (r, v) in zip(mixed, alone):
        bitwise(j.result[0], r)
        bitwise(j.result[1], v)
        assert j.latency > 0


def test_autotuner_picks_a_candidate_and_results_do_not_depend_on_it(maps, monkeypatch):
    from pypatools_amd import autotune
    grid, f1, f2, fields = maps
    monkeypatch.delenv("PYPAT_AUTOTUNE_CACHE", raising=False)
    rng = np.random.default_rng(35)
    bunches = _bunches(rng, [5000, 700, 3000, 1500], ["boris", "vay", "rk4", "boris"])
    cfg = autotune.autotune_batch(fields, bunches, "fp32_mixed", sample_steps=120, use_cache=True)
    assert cfg["block_size"] in autotune.BLOCKS and cfg["steps_per_launch"] in autotune.STEPS_PER_LAUNCH
    assert cfg["sort_every"] in autotune.SORT_EVERY
    # the second call comes from the per-process cache (no file under $HOME)
    t = __import__("time").perf_counter()
    assert autotune.autotune_batch(fields, bunches, "fp32_mixed", sample_steps=120) == cfg
    assert __import__("time").perf_counter() - t < 0.05
    tuned = BatchedTracker(fields, "fp32_mixed", **cfg).load(bunches)
    tuned.run()
    default = BatchedTracker(fields, "fp32_mixed", **autotune.DEFAULT).load(bunches)
    default.run()
    for (r, v), (rr, vv) in zip(tuned.results(), default.results()):
        bitwise(r, rr)
        bitwise(v, vv)
