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
VALU_PEAK_TFLOPS = 78.6       # MI355X vector peak without packed FP32 (= FP64 vector): 256 CUs x 128 flop/clk x 2.4 GHz
VALU_ISSUE_PER_S = 256 * 4 * 2.4e9 / 4        # wave64 VALU instructions per second, whole chip
# sources of the hot kernel; counters in profiles/pmc_counters.json are used only for this exact source
def kernel_source_digest():
    """sha256 of the fast-path kernel sources the loaded library was built from (compiled into it by
    pypatools_amd/build.py); counters are matched to the binary that runs, not to the files on disk."""
    from pypatools_amd import _lib
    return _lib.kernel_digest()


def load_counters(precision):
    """Counters of k_hot for `precision` from profiles/pmc_counters.json, if they were taken on this source.
    Returns (record or None, status)."""
    cfile = os.path.join(ROOT, "profiles", "pmc_counters.json")
    if not os.path.exists(cfile):
        return None, "no profiles/pmc_counters.json"
    with open(cfile) as f:
        data = json.load(f)
    rec = data.get(precision)
    if rec is None:
        return None, f"no counters for {precision}"
    digest = kernel_source_digest()
    if rec.get("source_sha256") != digest:
        return None, f"stale: counters taken on source {str(rec.get('source_sha256'))[:12]}, kernel source is {digest[:12]}"
    return rec, "counters of this source"


def parse(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", default="fp32_mixed")
    ap.add_argument("--particles", type=int, default=1_000_000)
    ap.add_argument("--nsteps", type=int, default=10_000)
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="C2 line only: no loss, space-charge, C3/C4/C5 legs (profiling runs)")
    ap.add_argument("--steps-per-launch", type=int, default=0)
    ap.add_argument("--sort-every", type=int, default=500, help="steps between cell sorts (0: never)")
    ap.add_argument("--refine-bits", type=int, default=None,
                    help="sub-cell sort bins along the motion (2^bits); default: the engine's")
    ap.add_argument("--block", type=int, default=0, help="threads per block (0: library default 256)")
    ap.add_argument("--workload", default="c2", choices=["c1", "c2", "c3", "c4", "c5", "lossy", "spacecharge"])
    ap.add_argument("--bunches", type=int, default=512)
    ap.add_argument("--c4-scale", type=float, default=1.0, help="scale C4 bunch sizes (1.0 = the C4 config)")
    ap.add_argument("--autotune", action="store_true")
    ap.add_argument("--graph", action="store_true", help="replay the C4 window loop from a HIP graph")
    ap.add_argument("--jobs", type=int, default=300)
    ap.add_argument("--load", type=float, default=0.7, help="C5 offered load as a fraction of --capacity")
    ap.add_argument("--capacity", type=float, default=0.0,
                    help="C5 capacity (particle-steps/s) for lambda; 0: measure it on a closed batch of C5 jobs")
    ap.add_argument("--pool", type=int, default=60_000_000, help="C5 particle pool (HBM budget)")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run the AoS->SoA copy kernel over this many MiB (counter calibration) and exit")
    ap.add_argument("--summarize-pmc", action="store_true",
                    help="summarise rocprofv3 --pmc CSV output (--pmc-dirs, --copy-dirs) instead of running")
    ap.add_argument("--pmc-dirs", nargs="*", default=[])
    ap.add_argument("--copy-dirs", nargs="*", default=[])
    ap.add_argument("--kernel-regex", default=r"k_hot<1, true, true, false, false>")
    ap.add_argument("--pmc-steps-per-launch", type=int, default=500)
    ap.add_argument("--write-counters", action="store_true",
                    help="with --summarize-pmc: store the record in profiles/pmc_counters.json")
    args = ap.parse_args(argv)
    if args.refine_bits is None:
        from pypatools_amd.engine import DEFAULT_REFINE_BITS
        args.refine_bits = DEFAULT_REFINE_BITS
    return args


def read_pmc(dirs, kernel_regex):
    """Per-dispatch counter values of the kernels matching `kernel_regex` from rocprofv3 --pmc CSV output
    (*counter_collection.csv under each dir).  Returns {counter: mean per dispatch}, dispatch count,
    the matched kernel names and the mean dispatch duration in ns (when the CSV has timestamps)."""
    import csv
    import glob
    import re
    rx = re.compile(kernel_regex)
    per = {}            # (file, dispatch) -> {counter: value}
    dur = {}
    names = set()
    for d in dirs:
        for fn in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
            with open(fn, newline="") as f:
                for row in csv.DictReader(f):
                    name = row.get("Kernel_Name", "")
                    if not rx.search(name):
                        continue
                    names.add(name)
                    key = (fn, row.get("Dispatch_Id") or row.get("Correlation_Id"))
                    c = per.setdefault(key, {})
                    c[row["Counter_Name"]] = c.get(row["Counter_Name"], 0.0) + float(row["Counter_Value"])
                    if row.get("Start_Timestamp") and row.get("End_Timestamp"):
                        dur[key] = float(row["End_Timestamp"]) - float(row["Start_Timestamp"])
    if not per:
        return {}, 0, sorted(names), None
    counters = sorted({k for c in per.values() for k in c})
    mean = {k: float(np.mean([c[k] for c in per.values() if k in c])) for k in counters}
    return mean, len(per), sorted(names), (float(np.mean(list(dur.values()))) if dur else None)


def summarize_pmc(args):
    """Summaries of rocprofv3 --pmc runs: per-dispatch counter averages of the hot kernel, the HBM traffic
    scaled by the copy probe's factors, and (--write-counters) the record bench.py reads for `traffic` and
    the VALU issue fraction, stamped with the kernel source digest."""
    out = {"kernel_regex": args.kernel_regex, "source_sha256": kernel_source_digest()}
    res = {}
    for tag, dirs in (("kernel", args.pmc_dirs), ("copy", args.copy_dirs)):
        if not dirs:
            continue
        rx = args.kernel_regex if tag == "kernel" else r"k_util"
        mean, n, names, dur_ns = read_pmc(dirs, rx)
        res[tag] = mean
        out[tag] = {"dispatches": n, "kernels": names, "mean_per_dispatch": mean, "mean_duration_ns": dur_ns}
    k = res.get("kernel", {})
    c = res.get("copy", {})
    copy_bytes = args.copy_probe_mib * 1024 * 1024 // 24 * 24
    if c and k and "FETCH_SIZE" in k and "WRITE_SIZE" in k:
        f_fetch = c["FETCH_SIZE"] * 1024 / copy_bytes
        f_write = c["WRITE_SIZE"] * 1024 / copy_bytes
        hbm = (k["FETCH_SIZE"] / f_fetch + k["WRITE_SIZE"] / f_write) * 1024
        out["hbm"] = {"fetch_factor": f_fetch, "write_factor": f_write, "bytes_per_dispatch": hbm}
    units = args.particles * args.pmc_steps_per_launch
    waves = -(-args.particles // 64)
    if "SQ_INSTS_VALU" in k:
        out["per_wavefront_step"] = {name: k[name] / (waves * args.pmc_steps_per_launch)
                                     for name in ("SQ_INSTS_VALU", "SQ_INSTS_SALU", "SQ_INSTS_LDS", "SQ_INSTS_VMEM_RD",
                                                  "SQ_INSTS_SMEM") if name in k}
    dur_ns = out.get("kernel", {}).get("mean_duration_ns")
    if "GRBM_GUI_ACTIVE" in k and dur_ns:
        # GRBM_GUI_ACTIVE sums the busy cycles of every XCD: the shader clock the kernel ran at
        xcc = num_xcc(args.pmc_dirs)
        out["clock"] = {"GHz": k["GRBM_GUI_ACTIVE"] / xcc / dur_ns, "xcc": xcc,
                        "note": "GRBM_GUI_ACTIVE / XCDs / dispatch duration, under the profiler"}
    print(json.dumps(out, indent=1), flush=True)
    if args.write_counters:
        cfile = os.path.join(ROOT, "profiles", "pmc_counters.json")
        data = {}
        if os.path.exists(cfile):
            with open(cfile) as f:
                data = json.load(f)
        rec = data.get(args.precision, {})
        if rec.get("source_sha256") != out["source_sha256"]:
            rec = {}
        rec.update({"kernel": out["kernel"]["kernels"][0] if out.get("kernel", {}).get("kernels") else None,
                    "workload": f"C2 Boris {args.precision}, {args.particles} particles",
                    "particles": args.particles, "steps_per_launch": args.pmc_steps_per_launch,
                    "source_sha256": out["source_sha256"]})
        if "hbm" in out:
            rec["hbm_bytes_per_launch"] = out["hbm"]["bytes_per_dispatch"]
            rec["hbm_bytes_per_particle_step"] = out["hbm"]["bytes_per_dispatch"] / units
        if "clock" in out:
            rec["clock_GHz"] = out["clock"]["GHz"]
        for name, key in (("SQ_INSTS_VALU", "valu_insts_per_wavefront_step"),
                          ("SQ_INSTS_LDS", "lds_insts_per_wavefront_step"),
                          ("SQ_INSTS_VMEM_RD", "vmem_insts_per_wavefront_step"),
                          ("SQ_INSTS_SALU", "salu_insts_per_wavefront_step")):
            if name in out.get("per_wavefront_step", {}):
                rec[key] = out["per_wavefront_step"][name]
        data[args.precision] = rec
        with open(cfile, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)


def num_xcc(dirs):
    """XCD count of the profiled GPU (rocprofv3 agent info), 8 on MI355X when absent."""
    import csv
    import glob
    for d in dirs:
        for fn in glob.glob(os.path.join(d, "**", "*agent_info.csv"), recursive=True):
            with open(fn, newline="") as f:
                for row in csv.DictReader(f):
                    if row.get("Agent_Type") == "GPU" and row.get("Num_Xcc"):
                        return int(row["Num_Xcc"])
    return 8


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


def gather_probe(fe, fb, nthreads, iters):
    """Measured peak of the gather path: the hot kernel's 12 x 16 B buffer loads per cell, no push math.

    mode 0 = one cell per 64-lane wavefront (the vector memory path's return-data limit, the
    ceiling for cell-sorted particles), mode 1 = a cell per lane from an L2-resident footprint.
    Returns GB/s of gather bytes (192 B per cell read) for each, timed with HIP events.
    """
    from pypatools_amd import _lib
    from pypatools_amd.batch import make_map
    from pypatools_amd.device import DeviceArray, Event, synchronize
    m = make_map(fe, fb)
    out = DeviceArray((nthreads,), np.float32)
    res = {}
    for mode, ncell, name in ((0, 1024, "wave_uniform_cells_GBps"), (1, 4096, "per_lane_cells_L2_GBps")):
        _lib.call("pp_gather_probe", m.handle, mode, ncell, iters, nthreads, out.ptr, None)
        e0, e1 = Event(), Event()
        e0.record()
        _lib.call("pp_gather_probe", m.handle, mode, ncell, iters, nthreads, out.ptr, None)
        e1.record()
        synchronize()
        res[name] = nthreads * iters * 192 / (e0.elapsed_ms(e1) * 1e-3) / 1e9
    return res


def moments_bandwidth(state, reps=5):
    """Device beam moments (two reduction passes over the SoA state): time per call and HBM GB/s.

    Algorithmic bytes: 6 float64 per particle per pass, 2 passes.
    """
    from pypatools_amd import _lib
    from pypatools_amd.device import DeviceArray, Event, synchronize
    import ctypes
    sums, cen = DeviceArray((_lib.MOMENT_SUMS,)), DeviceArray((_lib.MOMENT_CENTRAL,))
    nb = ctypes.c_size_t()
    _lib.call("pp_beam_workspace_bytes", ctypes.byref(nb))
    ws = DeviceArray((nb.value,), np.uint8)
    args = state._state_args()

    def call():
        _lib.call("pp_beam_moments", args, 0, 938.272, 0, sums.ptr, cen.ptr, ws.ptr, ws.nbytes, None)

    call()
    e0, e1 = Event(), Event()
    e0.record()
    for _ in range(reps):
        call()
    e1.record()
    synchronize()
    t = e0.elapsed_ms(e1) * 1e-3 / reps
    return {"particles": state.n, "ms_per_call": t * 1e3, "GBps": 2 * 48 * state.n / t / 1e9,
            "hbm_frac": 2 * 48 * state.n / t / 1e9 / HBM_PEAK_GBS}


def collision_rate(st_old, st_new, reps=5):
    """Electrode loss test: every particle's step segment against a 128 x 64 beam-pipe mesh (16384 triangles)."""
    from pypatools_amd.device import DeviceArray, Event, synchronize
    from pypatools_amd.electrodes import TriangleElectrodes
    pipe = TriangleElectrodes.cylinder(0.06, 0.0, 0.128, n_phi=128, n_z=64)
    n = st_old.n
    hit = DeviceArray((n,), np.uint8)

    def call():
        pipe.collide_device((st_old.x.ptr, st_old.y.ptr, st_old.z.ptr), (st_new.x.ptr, st_new.y.ptr, st_new.z.ptr),
                            n, None, hit.ptr)

    call()
    e0, e1 = Event(), Event()
    e0.record()
    for _ in range(reps):
        call()
    e1.record()
    synchronize()
    t = e0.elapsed_ms(e1) * 1e-3 / reps
    return {"segments": n, "triangles": pipe.n_triangles, "ms_per_call": t * 1e3, "segments_per_s": n / t,
            "hits": int(hit.to_host().sum())}


def lossy_tracking(w, fp, qm, nsteps=2000, radii=(3e-3, 10e-3), precision="fp32_mixed"):
    """The C2 bunch inside a beam pipe along its whole path (64 x 256 quads): the loss test fused into
    the step loop (pp_push_lossy) with compaction of the live slots every 500 steps (partition + cell
    sort), and without compaction; against the same steps with no mesh, with an alive mask (the hot
    kernel's loss build, no test) and without.  A 3 mm pipe loses about 28 % of the bunch (sigma
    2 mm); a 10 mm pipe holds the bunch (the common case: the test runs, nothing is near a wall)."""
    from pypatools_amd.device import DeviceArray, synchronize
    from pypatools_amd.electrodes import TriangleElectrodes
    from pypatools_amd.engine import ParticleState, Tracker
    dt = w["dt"]
    M = len(w["r0"])
    length = float(np.max(w["v0"][:, 2])) * nsteps * dt + 0.2

    def run(mesh, sort_every, alive=True):
        st = ParticleState.from_host(w["r0"], w["v0"])
        if alive:
            st.alive = DeviceArray.from_host(np.ones(M, dtype=np.uint8))
        tr = Tracker("boris", precision, qm, sort_every=sort_every)
        active = np.ones(M, dtype=bool)
        synchronize()
        t = time.perf_counter()
        if mesh is not None:
            tr._lossy_loop(st, fp, nsteps, dt, None, None, mesh, active, None, False)
        else:
            tr._fused_loop(st, fp, nsteps, dt, 0, None, None, None)
        synchronize()
        el = time.perf_counter() - t
        out = {"s": el, "particle_steps_per_s": M * nsteps / el}
        if mesh is not None:
            out["lost_fraction"] = float(1 - active.mean())
        return out

    pipes = [TriangleElectrodes.cylinder(radius, 0.0, length, n_phi=64, n_z=256) for radius in radii]

    def measure():
        res = {"particles": M, "steps": nsteps, "precision": precision, "pipe_length_m": length,
               "no_mesh_alive_mask": run(None, 500), "no_mesh_no_mask": run(None, 500, alive=False),
               "pipes": []}
        for radius, pipe in zip(radii, pipes):
            e = {"pipe_radius_m": radius, "triangles": pipe.n_triangles, "compacted": run(pipe, 500)}
            if radius == radii[0]:
                e["not_compacted"] = run(pipe, 0)
            res["pipes"].append(e)
        return res

    measure()   # untimed pass: first launches load each kernel's code and make the first allocations
    return measure()


def space_charge_timing(state, cells=(128, 128, 128), extent=(0.128, 0.128, 0.128), reps=3, electrodes=False):
    """Device space-charge step of the resident bunch: CIC deposit, Poisson solve, E = -grad(phi).  Without
    electrodes the solve is direct (sine transforms); with electrodes (four grounded rods of 10 x 10 mm at
    30 mm from the axis, along the whole domain) it is BiCGSTAB on the Shortley-Weller system to a relative
    residual of 1e-10, preconditioned by the direct box solve."""
    from pypatools_amd.device import Event, synchronize
    from pypatools_amd.electrodes import TriangleElectrodes, box_mesh
    from pypatools_amd.space_charge import SpaceChargeSolver
    assy = None
    t_setup = 0.0
    if electrodes:
        rods = [box_mesh([cx - 0.005, cy - 0.005, -1.0], [cx + 0.005, cy + 0.005, 1.0])
                for cx, cy in ((0.03, 0.0), (-0.03, 0.0), (0.0, 0.03), (0.0, -0.03))]
        v = np.concatenate([r[0] for r in rods])
        t = np.concatenate([r[1] + 8 * k for k, r in enumerate(rods)])
        assy = TriangleElectrodes(v, t)
    t0 = time.perf_counter()
    s = SpaceChargeSolver(extent, cells, electrode_assembly=assy)
    t_setup = time.perf_counter() - t0
    q = 1e-9 / state.n                                  # a 1 nC bunch
    s._solve_device(state.x.ptr, state.y.ptr, state.z.ptr, None, q, state.n)
    e0, e1 = Event(), Event()
    e0.record()
    for _ in range(reps):
        s._solve_device(state.x.ptr, state.y.ptr, state.z.ptr, None, q, state.n)
    e1.record()
    synchronize()
    ms = e0.elapsed_ms(e1) / reps
    phi = s.phi.to_host()
    out = {"particles": state.n, "mesh": list(cells), "ms_per_solve": ms, "phi_max_V": float(phi.max())}
    if electrodes:
        it, res = s.solve_stats()
        out.update({"electrodes": "4 grounded rods, 48 triangles", "setup_s": t_setup, "iterations": it,
                    "relative_residual": res, "conductor_nodes": int((s.cell_type == 2).sum())})
    return out


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
    # the reference's as-shipped no-numba structure (pusher.py:755-760): the fields of all particles
    # through the vectorised gather, then the Boris kernel particle by particle on (3,) arrays
    from oracle import kernels as K
    Ms = 2000
    r2, v2 = w["r0"][:Ms].copy(), w["v0"][:Ms].copy()
    qm = OS.q_over_m("proton")
    steps2 = 0
    t1 = time.perf_counter()
    while time.perf_counter() - t1 < seconds / 3:
        E, B = fe(r2), fb(r2)
        for i in range(Ms):
            rn, vn = K.boris_step(r2[i:i + 1], v2[i:i + 1], E[i:i + 1], B[i:i + 1], w["dt"], qm, K.sq_pow)
            r2[i], v2[i] = rn[0], vn[0]
        steps2 += 1
    el2 = time.perf_counter() - t1
    return {"value": M * steps / el, "unit": "particle-steps/s", "cores": 1, "kind": "port",
            "sample": f"C2 Boris + trilinear, {M} protons x {steps} steps, oracle/ canonical NumPy FP64 "
                      f"(vectorised over particles, single thread), {el:.1f} s",
            "as_shipped": {"value": Ms * steps2 / el2, "unit": "particle-steps/s", "cores": 1,
                           "sample": f"{Ms} protons x {steps2} steps of the reference's no-numba structure "
                                     f"(vectorised gather, per-particle Boris kernel, libm pow squares), "
                                     f"{el2:.1f} s"},
            "host": host_info()}


def host_info():
    """The host the CPU baseline ran on: CPU model, os.cpu_count() and the process's affinity size."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        affinity = None
    return {"cpu_model": model, "cpu_count": os.cpu_count(), "affinity": affinity}


def run_c4(args, reps=None, warmup=None):
    """C4: 512 bunches (~4e7 protons/H2+/electrons), one batched launch per window, autotuned.
    Returns (value, bunches, tracker, extra, cfg)."""
    import pypatools_amd as pa
    from pypatools_amd.autotune import autotune_batch
    from pypatools_amd.batch import BatchedTracker, c4_bunches
    from pypatools_amd.device import synchronize
    from pypatools_amd.workloads import c2_map
    grid, E, B = c2_map(128)
    fields = [(pa.Field.from_arrays(grid, E, periodic_z=True), pa.Field.from_arrays(grid, B, periodic_z=True))]
    t = time.perf_counter()
    bunches = c4_bunches(nbunch=args.bunches, scale=args.c4_scale)
    gen_s = time.perf_counter() - t
    tune = {}
    if args.autotune:
        sub = bunches[::8]
        t = time.perf_counter()
        cfg = autotune_batch(fields, sub, args.precision, sample_steps=1000, use_cache=False)
        tune = {"autotune_s": time.perf_counter() - t, "config": cfg}
    else:
        cfg = {"block_size": 256, "steps_per_launch": 500, "sort_every": 500}
    bt = BatchedTracker(fields, args.precision, use_graph=args.graph, **cfg)
    rates = []
    reps = args.steps if reps is None else reps
    warmup = args.warmup if warmup is None else warmup
    for rep in range(warmup + reps):
        bt.load(bunches)
        synchronize()
        t = time.perf_counter()
        bt.run()
        el = time.perf_counter() - t
        if rep >= warmup:
            rates.append(bt.particle_steps / el)
    res = bt.results()
    r_all = np.concatenate([r for r, _ in res])
    inside = float(np.mean((np.abs(r_all[:, 0]) <= 0.0635) & (np.abs(r_all[:, 1]) <= 0.0635)))
    value = float(np.median(rates))
    extra = {"generate_s": gen_s, "in_map_fraction_end": inside, "rates": rates, **tune}
    return value, bunches, bt, extra, cfg


def bench_c4(args):
    value, bunches, bt, extra, cfg = run_c4(args)
    out = {"metric": "particle-steps/sec on 1 MI355X (C4 batched persistent launch)", "value": value,
           "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": 1e3 * bt.particle_steps / value, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32+f64" if args.precision == "fp32_mixed" else "f64",
           "data": "synthetic",
           "config": {"workload": f"C4: {len(bunches)} bunches, {sum(b.n for b in bunches)} particles, "
                                  f"{bt.particle_steps:.3e} particle-steps, shared 128^3 map, per-bunch scale "
                                  "0.5-1.5, species proton/H2+/electron, Boris or Vay, 1e3-1e4 steps",
                      "precision": args.precision, "hip_graph": bool(args.graph), **cfg},
           "extra": extra}
    print(json.dumps(out), flush=True)


def _gamma(v):
    from pypatools_amd.constants import CLIGHT
    b2 = np.sum(v * v, axis=1) / CLIGHT ** 2
    return 1.0 / np.sqrt(1.0 - b2)


def _track_rate(tracker, st, fp, nsteps, dt, reps=1, t0=0.0):
    """Median particle-steps/s of `reps` runs of nsteps from the resident initial state `st`."""
    from pypatools_amd.device import synchronize
    from pypatools_amd.engine import ParticleState
    rates, last = [], None
    for _ in range(reps):
        s = ParticleState(st.n)
        for d, q in zip((s.x, s.y, s.z, s.vx, s.vy, s.vz), (st.x, st.y, st.z, st.vx, st.vy, st.vz)):
            d.copy_from(q)
        synchronize()
        t = time.perf_counter()
        tracker.run_resident(s, fp, nsteps, dt, step0=0, rf=tracker._rf(fp, t0, dt, nsteps))
        synchronize()
        rates.append(st.n * nsteps / (time.perf_counter() - t))
        last = s
    last.unsort()
    return float(np.median(rates)), last.to_host()


def run_c3(args, reps=None):
    """C3: RK4_rel through a TM010-like RF map (E cos, B sin), 4e6 electrons, 5000 steps; FP32-mixed vs FP64.
    Returns (value, extra, M, N, phase)."""
    import pypatools_amd as pa
    from pypatools_amd.engine import FieldPair, ParticleState, Tracker
    from pypatools_amd.workloads import c3_bunch, c3_map
    M = args.particles if args.particles != 1_000_000 else 4_000_000
    grid, E, B, w = c3_map()
    r0, v0, phase = c3_bunch(M)
    N, dt = 5000, 2e-13
    fe = pa.RFField(pa.Field.from_arrays(grid, E), w, phase, "cos")
    fb = pa.RFField(pa.Field.from_arrays(grid, B), w, phase, "sin")
    fp = FieldPair(fe, fb)
    qm = pa.IonSpecies("electron").q_over_m
    st0 = ParticleState.from_host(r0, v0)
    res = {}
    for prec in ([args.precision] + ([] if args.no_fp64 else ["fp64"])):
        tr = Tracker("rk4_rel", prec, qm, sort_every=args.sort_every, refine_bits=args.refine_bits)
        res[prec] = _track_rate(tr, st0, fp, N, dt, reps=(args.steps if reps is None else reps)
                                if prec == args.precision else 1)
    value, (rm, vm) = res[args.precision]
    extra = {}
    if "fp64" in res:
        r64, v64 = res["fp64"][1]
        path = np.linalg.norm(v0, axis=1) * N * dt
        dr = np.linalg.norm(rm - r64, axis=1)
        g64, gm = _gamma(v64), _gamma(vm)
        extra = {"fp64_particle_steps_per_s": res["fp64"][0], "pos_err_rel_path_max": float(np.max(dr / path)),
                 "pos_err_max_m": float(np.max(dr)), "energy_err_rel_max": float(np.max(np.abs(gm - g64) / (g64 - 1))),
                 "speedup_over_fp64": value / res["fp64"][0]}
    extra["mean_kinetic_gain_keV"] = float(np.mean(_gamma(vm) - _gamma(v0)) * 510.99895069)
    # inside the map at the end: |x|, |y| within the transverse half-width and z within [z0, z_last]
    half = float(grid["x"][-1])
    extra["in_map_fraction_end"] = float(np.mean((np.abs(rm[:, 0]) <= half) & (np.abs(rm[:, 1]) <= half)
                                                 & (rm[:, 2] >= grid["z"][0]) & (rm[:, 2] <= grid["z"][-1])))
    extra["z_end_range_m"] = [float(rm[:, 2].min()), float(rm[:, 2].max())]
    return value, extra, M, N, phase


def bench_c3(args):
    value, extra, M, N, phase = run_c3(args)
    out = {"metric": "particle-steps/sec on 1 MI355X (C3 RK4_rel through an RF map)", "value": value,
           "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": 0,
           "ms_per_step": 1e3 * M * N / value, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32+f64" if args.precision == "fp32_mixed" else "f64", "data": "synthetic",
           "config": {"workload": f"C3: 96x96x256 TM010-like map (E0 10 MV/m, 100 MHz, E cos / B sin, phase "
                                  f"{phase:.3f}), {M} electrons at 100 keV in a 3 mm disk, dt 2e-13 s, {N} steps, "
                                  "rk4_rel", "precision": args.precision, "particles": M},
           "extra": extra}
    print(json.dumps(out), flush=True)


def c2_algorithms(w, fp, qm, args):
    """The C2 track with the other pushers of the C2 config (textbook vay, the reference's vay_rel): the rate in
    FP32-mixed and strict FP64, and the FP32-mixed position error against strict FP64 (= the oracle)."""
    from pypatools_amd.device import synchronize
    from pypatools_amd.engine import U_STATE, ParticleState, Tracker
    M, N, dt = len(w["r0"]), args.nsteps, w["dt"]
    path = np.linalg.norm(w["v0"], axis=1) * N * dt
    out = {}
    for alg in ("vay", "vay_rel"):
        res, pos = {}, {}
        for prec in ("fp32_mixed", "fp64"):
            st = ParticleState.from_host(w["r0"], w["v0"])
            if alg in U_STATE:
                st.v_to_u()
            tr = Tracker(alg, prec, qm, block_size=args.block, sort_every=args.sort_every,
                         refine_bits=args.refine_bits, refine_axis=2)
            synchronize()
            t = time.perf_counter()
            tr.run_resident(st, fp, N, dt)
            synchronize()
            res[f"{prec}_particle_steps_per_s"] = M * N / (time.perf_counter() - t)
            st.unsort()
            pos[prec] = st.to_host()[0]
            del st
        res["pos_err_rel_path_max"] = float(np.max(np.linalg.norm(pos["fp32_mixed"] - pos["fp64"], axis=1) / path))
        out[alg] = res
    return out


def bench_c1(args):
    """C1: FP64 Boris in a uniform 1 T field, 1e4 protons x 1000 steps; error vs the closed-form gyro-orbit."""
    import pypatools_amd as pa
    from pypatools_amd.engine import FieldPair, ParticleState, Tracker
    from pypatools_amd.workloads import c1
    w = c1()
    r0, v0, N, dt = w["r0"], w["v0"], w["nsteps"], w["dt"]
    fp = FieldPair(pa.Field(dim=0, field=dict(zip("xyz", w["E"]))), pa.Field(dim=0, field=dict(zip("xyz", w["B"]))))
    qm = pa.IonSpecies("proton").q_over_m
    st0 = ParticleState.from_host(r0, v0)
    value, (r, v) = _track_rate(Tracker("boris", "fp64", qm), st0, fp, N, dt, reps=args.steps)
    # closed form of the staggered Boris map in uniform B along z: v rotates by -theta per step,
    # theta = 2 atan(qm B dt / 2); r_N = r_0 + dt sum_{n=1..N} R^n v_0
    th = 2.0 * np.arctan(qm * 1.0 * dt / 2.0)
    n = np.arange(1, N + 1)
    c, s = np.cos(n * th).sum(), np.sin(n * th).sum()
    rx = r0[:, 0] + dt * (c * v0[:, 0] + s * v0[:, 1])
    ry = r0[:, 1] + dt * (-s * v0[:, 0] + c * v0[:, 1])
    rz = r0[:, 2] + dt * N * v0[:, 2]
    exact = np.column_stack([rx, ry, rz])
    path = np.linalg.norm(v0, axis=1) * N * dt
    err = float(np.max(np.linalg.norm(r - exact, axis=1) / path))
    # boris_rel: the state holds u = gamma v; in pure B gamma is constant and u rotates by
    # theta_rel = 2 atan(qm B dt / (2 gamma)) per step, the position advancing by dt u / gamma
    from pypatools_amd.constants import CLIGHT
    st_rel = ParticleState.from_host(r0, v0)
    st_rel.v_to_u()
    rate_rel, (r_rel, _) = _track_rate(Tracker("boris_rel", "fp64", qm), st_rel, fp, N, dt, reps=args.steps)
    gam = 1.0 / np.sqrt(1.0 - np.sum(v0 * v0, axis=1) / CLIGHT ** 2)
    th_r = 2.0 * np.arctan(qm * 1.0 * dt / (2.0 * gam))
    nn = np.arange(1, N + 1)[None, :]
    c_r, s_r = np.cos(nn * th_r[:, None]).sum(axis=1), np.sin(nn * th_r[:, None]).sum(axis=1)
    exact_rel = np.column_stack([r0[:, 0] + dt * (c_r * v0[:, 0] + s_r * v0[:, 1]),
                                 r0[:, 1] + dt * (-s_r * v0[:, 0] + c_r * v0[:, 1]), r0[:, 2] + dt * N * v0[:, 2]])
    err_rel = float(np.max(np.linalg.norm(r_rel - exact_rel, axis=1) / path))
    larmor = float(np.mean(np.linalg.norm(v0[:, :2], axis=1) / (qm * 1.0)))
    out = {"metric": "particle-steps/sec on 1 MI355X (C1 plain FP64 Boris, uniform B)", "value": value,
           "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": 0,
           "ms_per_step": 1e3 * len(r0) * N / value, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "C1: 1e4 protons at 1 MeV, isotropic, uniform B = 1 T along z, E = 0, "
                                  "dt 1e-11 s, 1000 steps, boris (FP64)", "precision": "fp64"},
           "extra": {"pos_err_rel_path_max_vs_closed_form": err, "theta_per_step": th,
                     "boris_rel": {"particle_steps_per_s": rate_rel, "pos_err_rel_path_max_vs_closed_form": err_rel,
                                   "theta_per_step_mean": float(np.mean(th_r))},
                     "mean_transverse_larmor_radius_m": larmor}}
    print(json.dumps(out), flush=True)


def run_c5(args, n_jobs=None):
    """C5: Poisson arrivals, continuous batching, p50/p99 latency at a load set from measured capacity.
    Returns (jobs, lam, capacity, measured, span, srv)."""
    import pypatools_amd as pa
    from pypatools_amd.device import synchronize
    from pypatools_amd.serving import TrackingServer, c5_jobs
    from pypatools_amd.workloads import c2_map
    grid, E, B = c2_map(128)
    fields = []
    for k in range(4):
        s = 1.0 + 0.1 * k
        Ek = {c: s * E[c] for c in "xyz"}
        Bk = {c: (2.0 - s) * B[c] for c in "xyz"}
        fields.append((pa.Field.from_arrays(grid, Ek, periodic_z=True), pa.Field.from_arrays(grid, Bk, periodic_z=True)))
    srv = TrackingServer(fields, args.precision, chunk_steps=128, pool_particles=args.pool)
    n_jobs = args.jobs if n_jobs is None else n_jobs
    # the job set of the run: its contents do not depend on the arrival rate (c5_jobs draws them in the
    # same order for any rate), so the closed batch below and the timed run push the same jobs
    cal = c5_jobs(n_jobs, 1e9)
    mean_work = float(np.mean([j.n * j.nsteps for j in cal]))
    capacity = args.capacity
    measured = None
    if capacity <= 0:
        # capacity of the server on this job set (Boris / Vay / RK4 jobs, RK4 costing 4 gathers per step;
        # a few long jobs hold most of the work, so the mix of one set differs from the next): the jobs
        # submitted at once as a closed batch, timed until the queue has emptied and fewer than 8 jobs are
        # left running (the drain of the last few long jobs is not capacity)
        warm = c5_jobs(4, 1e9, seed=12)          # first admissions allocate the page-locked buffers
        for j in warm:
            srv.submit(j)
        srv.run_until_idle()
        srv.particle_steps = 0
        t = time.perf_counter()
        for j in cal:
            srv.submit(j)
        el, done = None, 0
        while not srv.idle:
            srv.round(time.perf_counter())
            if el is None and not srv.queue and len(srv.running) < 8:
                synchronize()
                el = time.perf_counter() - t
                done = srv.particle_steps          # every enqueued chunk has completed
        measured = done / el
        capacity = measured
        srv.particle_steps = 0
        srv.rounds = 0
    lam = args.load * capacity / mean_work
    jobs = c5_jobs(n_jobs, lam)
    t0 = time.perf_counter()
    for j in jobs:
        j.arrival += t0
    pending = list(jobs)
    while pending or not srv.idle:
        now = time.perf_counter()
        while pending and pending[0].arrival <= now:
            srv.submit(pending.pop(0))
        if srv.idle and pending:
            time.sleep(max(0.0, min(pending[0].arrival - time.perf_counter(), 0.05)))
            continue
        srv.round(now)
    span = time.perf_counter() - t0
    return jobs, lam, capacity, measured, span, srv


def c5_summary(jobs, lam, capacity, measured, span, srv, load):
    lat = np.array([j.latency for j in jobs])
    work = sum(j.n * j.nsteps for j in jobs)
    return {"jobs": len(jobs), "offered_load": load, "arrival_rate_jobs_per_s": lam,
            "capacity_particle_steps_per_s": capacity, "capacity_measured": measured is not None,
            "p50_s": float(np.percentile(lat, 50)), "p90_s": float(np.percentile(lat, 90)),
            "p99_s": float(np.percentile(lat, 99)), "mean_s": float(lat.mean()),
            "throughput_particle_steps_per_s": work / span, "makespan_s": span, "rounds": srv.rounds,
            "pool_particles": srv.pool_n}


def bench_c5(args):
    jobs, lam, capacity, measured, span, srv = run_c5(args)
    summ = c5_summary(jobs, lam, capacity, measured, span, srv, args.load)
    out = {"metric": "p99 job latency, continuous-batching tracking server on 1 MI355X (C5)",
           "value": summ["p99_s"], "unit": "s", "n_gpus": 1, "steps": 1, "warmup": 0,
           "ms_per_step": span * 1e3, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32+f64" if args.precision == "fp32_mixed" else "f64", "data": "synthetic",
           "config": {"workload": f"C5: {len(jobs)} jobs, Poisson arrivals at {lam:.2f} jobs/s "
                                  f"({args.load:.0%} of {capacity:.2e} particle-steps/s "
                                  f"{'measured ' if measured else ''}capacity), N log-uniform "
                                  "1e3-2e6, steps log-uniform 1e2-2e4, Boris/Vay/RK4, 4 resident 128^3 maps, "
                                  "128-step chunks", "precision": args.precision, "pool_particles": srv.pool_n},
           "extra": summ}
    print(json.dumps(out), flush=True)


def bench_lossy(args):
    """C2 bunch and map with an electrode mesh: track_batch's loss test on the device (extension row)."""
    import pypatools_amd as pa
    from pypatools_amd.engine import FieldPair
    from pypatools_amd.workloads import c2
    w = c2(M=args.particles)
    fp = FieldPair(pa.Field.from_arrays(w["grid"], w["E"], periodic_z=True),
                   pa.Field.from_arrays(w["grid"], w["B"], periodic_z=True))
    qm = pa.IonSpecies("proton").q_over_m
    t0 = time.perf_counter()
    res = lossy_tracking(w, fp, qm, precision=args.precision)
    el = time.perf_counter() - t0
    main_pipe = res["pipes"][0]["compacted"]
    out = {"metric": "particle-steps/s of Boris tracking with the electrode loss test (C2 bunch in a beam pipe)",
           "value": main_pipe["particle_steps_per_s"], "unit": "particle-steps/s", "n_gpus": 1, "steps": 1,
           "warmup": 0, "ms_per_step": 1e3 * main_pipe["s"], "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32+f64" if args.precision == "fp32_mixed" else "f64",
           "data": "synthetic",
           "config": {"workload": f"C2 bunch ({res['particles']} protons) and 128^3 map, {res['steps']} steps, "
                                  "cylindrical pipe meshes of 32768 triangles along the whole path; fused loss test, "
                                  "compaction every 500 steps", "precision": args.precision},
           "extra": dict(res, wall_s=el)}
    print(json.dumps(out), flush=True)


def bench_spacecharge(args):
    """Space-charge step of the C2 bunch (1e6 protons, 1 nC) on a 128^3 mesh, box domain and four grounded rods."""
    from pypatools_amd.engine import ParticleState
    from pypatools_amd.workloads import c2
    w = c2(M=args.particles)
    st = ParticleState.from_host(w["r0"] - np.array([0.0, 0.0, 0.0635]), w["v0"])     # bunch centred in the box
    box = space_charge_timing(st, reps=5)
    rods = space_charge_timing(st, reps=3, electrodes=True)
    out = {"metric": "ms per space-charge step (CIC deposit + Poisson solve + E = -grad phi), C2 bunch, 128^3 mesh",
           "value": box["ms_per_solve"], "unit": "ms", "n_gpus": 1, "steps": 5, "warmup": 1,
           "ms_per_step": box["ms_per_solve"], "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
           "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"C2 bunch ({st.n} protons, 1 nC) centred in a 0.128 m box, 128^3 cell-centre mesh; "
                                  "box domain (direct sine-transform solve) and four grounded rods (BiCGSTAB)"},
           "extra": {"box": box, "four_rods": rods}}
    print(json.dumps(out), flush=True)


def main():
    args = parse()
    if args.summarize_pmc:
        if not args.copy_probe_mib:
            args.copy_probe_mib = 768
        summarize_pmc(args)
        return
    if args.copy_probe_mib:
        copy_probe(args.copy_probe_mib)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the engine shards by particle "
              "index across GPUs by design but this build measures one GPU's share", file=sys.stderr)
        sys.exit(2)
    if args.workload == "c1":
        return bench_c1(args)
    if args.workload == "c3":
        return bench_c3(args)
    if args.workload == "c4":
        return bench_c4(args)
    if args.workload == "c5":
        return bench_c5(args)
    if args.workload == "lossy":
        return bench_lossy(args)
    if args.workload == "spacecharge":
        return bench_spacecharge(args)

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
        tr = Tracker("boris", prec, qm, block_size=args.block, sort_every=args.sort_every,
                     refine_bits=args.refine_bits, refine_axis=2)
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
                extra["pos_err_denominator"] = "|v0| N dt (path length); FP64 strict result = oracle bit for bit"
            del s64
        extra["c2_algorithms"] = c2_algorithms(w, fp, qm, args)

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
    for prec in ("fp64_fast", "fp64"):
        if f"{prec}_particle_steps_per_s" in extra:
            extra[f"speedup_over_{prec}"] = value / extra[f"{prec}_particle_steps_per_s"]
    # roofline of the dominant kernel (k_hot<boris>): the path has no dense contraction and its gathers are
    # served from LDS/L2 (HBM sees a fraction of a byte per particle-step), so the bound is the vector ALU:
    # algorithmic flops per launch (190 per particle-step) / launch time, against the 78.6 TFLOP/s vector peak
    units = M * steps_per_launch
    avg_s = avg_ms * 1e-3
    tflops = FLOPS_PER_STEP * units / avg_s / 1e12
    gather_gbs = GATHER_BYTES[args.precision] * units / avg_s / 1e9
    probe = gather_probe(fe, fb, M, steps_per_launch)
    gpeak = probe["wave_uniform_cells_GBps"]
    # counters of the same kernel and workload from the committed rocprofv3 runs on this source: HBM bytes
    # per launch (FETCH_SIZE + WRITE_SIZE, copy-calibrated) and VALU instructions per wavefront-step
    rec, cstatus = load_counters(args.precision)
    traffic, issue = None, {"status": cstatus}
    if rec is not None:
        if rec["steps_per_launch"] == steps_per_launch and rec.get("particles") == M:
            traffic = rec.get("hbm_bytes_per_launch")       # recorded for FP32-mixed only
        wave_steps = -(-M // 64) * steps_per_launch
        slots = VALU_ISSUE_PER_S * avg_s                     # wave64 VALU issue slots in one launch
        issue.update({"valu_insts_per_wavefront_step": rec["valu_insts_per_wavefront_step"],
                      "frac": rec["valu_insts_per_wavefront_step"] * wave_steps / slots,
                      "slots": "256 CUs x 4 SIMDs, one wave64 VALU instruction per 4 cycles at 2.4 GHz",
                      "flop_efficiency_per_instruction": FLOPS_PER_STEP / (2 * rec["valu_insts_per_wavefront_step"])})
        if rec.get("clock_GHz"):
            # the same at the shader clock the kernel ran at under the profiler (below the 2.4 GHz peak)
            issue["clock_GHz_measured"] = rec["clock_GHz"]
            issue["frac_at_measured_clock"] = issue["frac"] * 2.4 / rec["clock_GHz"]
    out = {
        "metric": "particle-steps/sec on 1 MI355X, FP32-mixed vs FP64 at stated error; p99 job latency",
        "value": value, "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * el / args.steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64" if args.precision == "fp32_mixed" else "f64", "data": "synthetic",
        "config": {"workload": "C2: Boris + trilinear gather, 128^3 E/B map (1 mm, periodic z), "
                               f"{M} protons 2.5 MeV Gaussian sigma 2 mm, dt 5e-11 s, {N} steps per bench step",
                   "precision": args.precision, "particles": M, "time_steps_per_bench_step": N},
        "roofline": {"bound": "valu", "achieved": tflops, "peak": VALU_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": tflops / VALU_PEAK_TFLOPS, "traffic": traffic,
                     "kernel": f"pp_fast::k_hot<1, {str(args.precision == 'fp32_mixed').lower()}, true, false, false> (boris, "
                               f"periodic z; {steps_per_launch} steps/launch, avg {avg_ms:.3f} ms by HIP events)",
                     "flops_per_particle_step": FLOPS_PER_STEP, "particle_steps_per_launch": units,
                     "peak_note": "vector peak without packed FP32 (not used on this path) = FP64 vector peak",
                     "valu_issue": issue,
                     "hbm": {"bytes_per_launch": traffic,
                             "bytes_per_particle_step": traffic / units if traffic else None,
                             "GBps": traffic / avg_s / 1e9 if traffic else None,
                             "frac": traffic / avg_s / 1e9 / HBM_PEAK_GBS if traffic else None},
                     "gather": {"algorithmic_bytes_per_particle_step": GATHER_BYTES[args.precision],
                                "GBps": gather_gbs, "vector_path_probe_peak_GBps": gpeak,
                                "frac_of_probe": gather_gbs / gpeak, "probe": probe,
                                "note": "cells are staged through LDS, so the gather rate can exceed the probe"}},
        "extra": extra,
    }
    # beam diagnostics on the resident state: the C2 bunch, and a 4e7-particle state (HBM-resident)
    big = ParticleState(40_000_000)
    for arr, val in zip((big.x, big.y, big.z, big.vx, big.vy, big.vz), (1e-3, -2e-3, 0.05, 1e4, -2e4, 2.2e7)):
        arr.fill(val)
    out["extra"]["beam_moments"] = [moments_bandwidth(st0), moments_bandwidth(big)]
    del big
    if not args.no_extras:
        out["extra"]["electrode_collision"] = collision_rate(st0, st)
        out["extra"]["lossy_tracking"] = lossy_tracking(w, fp, qm)
        out["extra"]["space_charge"] = space_charge_timing(st0)
        out["extra"]["space_charge_electrodes"] = space_charge_timing(st0, electrodes=True)
        del st, st0
        # the other workloads of the metric: C3 (RK4_rel + RF), C4 (batched) and C5 (served, p99)
        t = time.perf_counter()
        c3v, c3x, c3m, c3n, _ = run_c3(args, reps=1)
        out["extra"]["c3"] = dict(c3x, particle_steps_per_s=c3v, particles=c3m, steps=c3n, wall_s=time.perf_counter() - t)
        t = time.perf_counter()
        c4v, c4b, c4t, c4x, c4cfg = run_c4(args, reps=1, warmup=0)
        out["extra"]["c4"] = {"particle_steps_per_s": c4v, "bunches": len(c4b), "particles": int(sum(b.n for b in c4b)),
                              "particle_steps": c4t.particle_steps, "in_map_fraction_end": c4x["in_map_fraction_end"],
                              "config": c4cfg, "wall_s": time.perf_counter() - t}
        del c4t, c4b
        t = time.perf_counter()
        out["extra"]["c5"] = dict(c5_summary(*run_c5(args), args.load), wall_s=time.perf_counter() - t)
    out["config"]["sort_every"] = args.sort_every
    out["config"]["refine_bits"] = args.refine_bits
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(w)
        out["extra"]["gpu_over_cpu"] = value / out["cpu_baseline"]["value"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
