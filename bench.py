#!/usr/bin/env python3
"""Benchmark: particle-steps/s of Boris + trilinear gather on one MI355X (C2 workload).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision fp32_mixed]

One benchmark step = the full C2 track: 1e6 protons x 1e4 time steps
(1e10 particle-steps) of 'boris' through the 128^3 E/B map (periodic in z),
restarted from the initial bunch, which stays resident in HBM (the restart
is a device-to-device copy inside the timed region).  The C2 optics are
transversely unstable (the E quadrupole defocuses x): over the first 1e4 steps
every particle stays inside the map, later the beam leaves it, so each bench
step covers exactly the configured 1e4 steps.  Prints ONE JSON line.

Also measured in the same run (reported under "extra"): the FP64 paths on the
same segment, the FP32-mixed position error against the FP64 path (which the
parity tests show equals the FP64 oracle bit for bit) after 1e4 steps, and the
CPU oracle timed on a bounded sample on this host.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

GATHER_BYTES = {"fp32_mixed": 192, "fp64_fast": 384, "fp64": 384}   # 8 corners x 6 comps x sizeof(map element)
FLOPS_PER_STEP = 55 + 135                                          # non-rel Boris push + one 6-comp gather
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", default="fp32_mixed")
    ap.add_argument("--particles", type=int, default=1_000_000)
    ap.add_argument("--nsteps", type=int, default=10_000)
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--steps-per-launch", type=int, default=0)
    ap.add_argument("--sort-every", type=int, default=2000, help="steps between cell sorts (0: never)")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run the AoS->SoA copy kernel over this many MiB (counter calibration) and exit")
    return ap.parse_args()


def copy_probe(mib):
    """Run pp_aos_to_soa once over `mib` MiB of float64 (reads mib MiB, writes mib MiB)."""
    from pypatools_amd import _lib
    from pypatools_amd.device import DeviceArray, synchronize
    n = mib * 1024 * 1024 // 24
    src = DeviceArray((n, 3))
    src.fill(1.0)
    dst = [DeviceArray((n,)) for _ in range(3)]
    synchronize()
    _lib.call("pp_aos_to_soa", src.ptr, n, dst[0].ptr, dst[1].ptr, dst[2].ptr, None)
    synchronize()
    print(json.dumps({"copy_probe_bytes_read": n * 24, "copy_probe_bytes_written": n * 24}), flush=True)


def cpu_baseline(w, seconds=15.0):
    """The FP64 oracle (canonical vectorised NumPy form, one core) on a sample of C2."""
    from oracle.field import OPeriodicZ
    from oracle.pusher import OraclePusher
    from oracle import species as OS
    M = 20_000
    fe, fb = OPeriodicZ(w["grid"], w["E"]), OPeriodicZ(w["grid"], w["B"])
    pu = OraclePusher(OS.q_over_m("proton"), "boris")
    r, v = w["r0"][:M].copy(), w["v0"][:M].copy()
    steps = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        r, v = pu.push_batch(r, v, fe, fb, w["dt"])
        steps += 1
    el = time.perf_counter() - t0
    return {"value": M * steps / el, "unit": "particle-steps/s", "cores": 1, "kind": "port",
            "sample": f"C2 Boris + trilinear, {M} protons x {steps} steps, oracle/ canonical NumPy FP64 "
                      f"(vectorised over particles, single thread), {el:.1f} s"}


def main():
    args = parse()
    if args.copy_probe_mib:
        copy_probe(args.copy_probe_mib)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the engine shards by particle "
              "index across GPUs by design but this build measures one GPU's share", file=sys.stderr)
        sys.exit(2)

    import pypatools_amd as pa
    from pypatools_amd.device import DeviceArray, Event, synchronize
    from pypatools_amd.engine import FieldPair, ParticleState, Tracker
    from pypatools_amd.workloads import c2

    w = c2(M=args.particles)
    M, N, dt = args.particles, args.nsteps, w["dt"]
    qm = pa.IonSpecies("proton").q_over_m
    fe = pa.Field.from_arrays(w["grid"], w["E"], periodic_z=True)
    fb = pa.Field.from_arrays(w["grid"], w["B"], periodic_z=True)
    fp = FieldPair(fe, fb)
    spl = args.steps_per_launch or None

    def segment(prec, state, step0, timer=None):
        tr = Tracker("boris", prec, qm, sort_every=args.sort_every)
        return tr.run_resident(state, fp, N, dt, step0=step0, steps_per_launch=spl, timer=timer)

    st0 = ParticleState.from_host(w["r0"], w["v0"])    # resident initial bunch
    st = ParticleState(M)

    def restart():
        for d, s in zip((st.x, st.y, st.z, st.vx, st.vy, st.vz), (st0.x, st0.y, st0.z, st0.vx, st0.vy, st0.vz)):
            d.copy_from(s)
        st.reset_order()

    # ---- accuracy: the track in FP32-mixed and in FP64 from the same bunch
    restart()
    synchronize()
    extra = {}
    t = time.perf_counter()
    segment(args.precision, st, 0)
    synchronize()
    extra["first_segment_s"] = time.perf_counter() - t
    st.unsort()
    r_mixed, _ = st.to_host()
    half = 0.0635
    extra["in_map_fraction_end"] = float(np.mean((np.abs(r_mixed[:, 0]) <= half) & (np.abs(r_mixed[:, 1]) <= half)))
    if not args.no_fp64:
        for prec in ("fp64_fast", "fp64"):
            s64 = ParticleState.from_host(w["r0"], w["v0"])
            synchronize()
            t = time.perf_counter()
            segment(prec, s64, 0)
            synchronize()
            el = time.perf_counter() - t
            extra[f"{prec}_particle_steps_per_s"] = M * N / el
            if prec == "fp64":
                s64.unsort()
                r64, _ = s64.to_host()
                path = np.linalg.norm(w["v0"], axis=1) * N * dt
                dr = np.linalg.norm(r_mixed - r64, axis=1)
                extra["pos_err_rel_path_max"] = float(np.max(dr / path))
                extra["pos_err_max_m"] = float(np.max(dr))
            del s64

    # ---- warmup + timed bench steps (each restarts the track; inputs resident in HBM)
    for _ in range(args.warmup):
        restart()
        segment(args.precision, st, 0)
    synchronize()
    durations = []

    def timer(launch):
        e0, e1 = Event(), Event()
        e0.record()
        launch()
        e1.record()
        durations.append((e0, e1))

    t0 = time.perf_counter()
    for _ in range(args.steps):
        restart()
        segment(args.precision, st, 0, timer)
    synchronize()
    el = time.perf_counter() - t0
    kern_ms = [a.elapsed_ms(b) for a, b in durations]
    launches_per_seg = len(kern_ms) // max(args.steps, 1)
    steps_per_launch = N // max(launches_per_seg, 1)
    avg_ms = float(np.mean(kern_ms))
    work = args.steps * M * N
    value = work / el
    # roofline of the dominant kernel (k_push_mixed<boris>): algorithmic gather bytes per launch / launch time
    units = M * steps_per_launch
    achieved = GATHER_BYTES[args.precision] * units / (avg_ms * 1e-3) / 1e9
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            traffic = json.load(f).get(args.precision)
    out = {
        "metric": "particle-steps/sec on 1 MI355X, FP32-mixed vs FP64 at stated error; p99 job latency",
        "value": value, "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * el / args.steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64" if args.precision == "fp32_mixed" else "f64", "data": "synthetic",
        "config": {"workload": "C2: Boris + trilinear gather, 128^3 E/B map (1 mm, periodic z), "
                               f"{M} protons 2.5 MeV Gaussian sigma 2 mm, dt 5e-11 s, {N} steps per bench step",
                   "precision": args.precision, "particles": M, "time_steps_per_bench_step": N},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": f"k_push_mixed<boris> ({steps_per_launch} steps/launch, avg {avg_ms:.2f} ms)",
                     "algorithmic_bytes_per_particle_step": GATHER_BYTES[args.precision],
                     "gflops": FLOPS_PER_STEP * units / (avg_ms * 1e-3) / 1e9},
        "extra": extra,
    }
    out["config"]["sort_every"] = args.sort_every
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(w)
        out["extra"]["gpu_over_cpu"] = value / out["cpu_baseline"]["value"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
