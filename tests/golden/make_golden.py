"""Write golden input/output vectors by running the reference PyPATools in THIS container.

Only functions that no closed form, invariant or library routine pins are
recorded here (the reference ships no golden vectors for this path):

* ``vay_push_single``  -- the reference's non-textbook relativistic Vay kernel
  (src/PyPATools/pusher.py:216-279), bitwise.
* ``Pusher.track_batch`` / ``track`` / ``push`` / ``push_batch`` bookkeeping for
  every algorithm (pusher.py:508-1012): Boris half steps, record layout,
  NaN/zero fill, last-record overwrite, active mask.
* ``Field.__call__`` dispatch and scaling order for 0-D/1-D/2-D/3-D fields,
  ScaledField, CompositeField and Field.zero (field.py:213-504, 595-690), with
  the 'fast' (CoordinateMapper) and 'scipy' backends.
* ``IonSpecies.q_over_m`` for every preset (species.py:7-163).

The reference is imported from /root/reference/src in a scratch working
directory (its import deletes and recreates ``./temp``) with HOME pointed at the
scratch directory.  Two modules it imports are absent from this image and are
replaced by empty placeholders: ``py_electrodes`` (used only as a type
annotation, pusher.py:33,469) and ``h5py`` (used only by file loaders,
field.py:13).  NumPy is the only array backend (numba is absent), so the
reference runs its no-numba paths.

Usage (from the repo root):  python tests/golden/make_golden.py
Outputs: tests/golden/*.npz (plain arrays, allow_pickle=False).
"""
import os
import sys
import tempfile
import types

import numpy as np

OUT = os.path.dirname(os.path.abspath(__file__))
REF_SRC = "/root/reference/src"


def _import_reference():
    scratch = tempfile.mkdtemp(prefix="pypat_golden_")
    os.chdir(scratch)
    os.environ["HOME"] = scratch
    pe = types.ModuleType("py_electrodes")
    pe2 = types.ModuleType("py_electrodes.py_electrodes")

    class PyElectrodeAssembly:  # annotation-only placeholder
        pass

    pe2.PyElectrodeAssembly = PyElectrodeAssembly
    pe.py_electrodes = pe2
    sys.modules.setdefault("py_electrodes", pe)
    sys.modules.setdefault("py_electrodes.py_electrodes", pe2)
    sys.modules.setdefault("h5py", types.ModuleType("h5py"))
    sys.path.insert(0, REF_SRC)
    import warnings
    warnings.simplefilter("ignore")
    from PyPATools import field, pusher, species  # noqa: E402
    return pusher, field, species


def c2_map(nx, ny, nz, dx, rng, noise=0.05):
    """C2-style analytic map (solenoid Bz with paraxial Br, quadrupole E) plus seeded noise."""
    x = (np.arange(nx) - (nx - 1) / 2.0) * dx
    y = (np.arange(ny) - (ny - 1) / 2.0) * dx
    z = np.arange(nz) * dx
    L = nz * dx
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    bz = 0.5 * (1.0 + 0.3 * np.cos(2 * np.pi * Z / L))
    dbz = -0.5 * 0.3 * (2 * np.pi / L) * np.sin(2 * np.pi * Z / L)
    vals = {
        "E": {"x": 2e6 * X, "y": -2e6 * Y, "z": np.zeros_like(X)},
        "B": {"x": -0.5 * X * dbz, "y": -0.5 * Y * dbz, "z": bz},
    }
    for fb, scale in (("E", 2e4), ("B", 0.05)):
        for c in "xyz":
            vals[fb][c] = vals[fb][c] + noise * scale * rng.standard_normal(X.shape)
    return {"x": x, "y": y, "z": z}, vals


class _WallAssembly:
    """Mock electrode assembly: particles crossing the plane x = wall are lost."""

    def __init__(self, wall):
        self.wall = wall

    def segment_intersects_surface(self, r_old, r_new):
        return {"hit_mask": (r_old[:, 0] <= self.wall) & (r_new[:, 0] > self.wall)}


def main():
    pusher, field, species = _import_reference()
    P, F, S = pusher.Pusher, field.Field, species.IonSpecies
    rng = np.random.default_rng(20260101)
    algs = list(P.ALGORITHMS)

    # ---- species ---------------------------------------------------------
    names = list(species.PRESETS.keys())
    np.savez(os.path.join(OUT, "species_qm.npz"),
             names=np.array(names),
             q_over_m=np.array([S(n).q_over_m for n in names]),
             mass_mev=np.array([S(n).mass_mev for n in names]),
             mass_kg=np.array([S(n).mass_kg for n in names]))

    # ---- vay_push_single --------------------------------------------------
    c = pusher.CLIGHT
    n = 256
    dirs = rng.standard_normal((n, 3))
    dirs /= np.linalg.norm(dirs, axis=1)[:, None]
    beta = np.concatenate([rng.uniform(0.0, 0.99, n - 8), [0.0, 0.5, 0.9999, 0.99995, 0.999999, 0.3, 0.7, 0.95]])
    v = dirs * (beta * c)[:, None]
    E = rng.standard_normal((n, 3)) * 10.0 ** rng.uniform(3, 8, (n, 1))
    B = rng.standard_normal((n, 3)) * 10.0 ** rng.uniform(-3, 1, (n, 1))
    E[:16] = 0.0
    B[16:32] = 0.0
    dt = 10.0 ** rng.uniform(-13, -9, n)
    dt[-4:] *= -1.0
    qm = np.where(rng.uniform(size=n) < 0.5, S("electron").q_over_m, S("proton").q_over_m)
    out = np.array([pusher.vay_push_single(v[i], E[i], B[i], dt[i], qm[i]) for i in range(n)])
    np.savez(os.path.join(OUT, "vay_single.npz"), v=v, E=E, B=B, dt=dt, qm=qm, out=out)

    # ---- Field.__call__ ---------------------------------------------------
    grid, vals = c2_map(12, 10, 14, 1e-3, rng)
    pts_in = np.column_stack([rng.uniform(g[0], g[-1], 3000) for g in (grid["x"], grid["y"], grid["z"])])
    # on faces / nodes / just outside
    faces = []
    for a, g in enumerate((grid["x"], grid["y"], grid["z"])):
        for val in (g[0], g[-1], g[-1] + 1e-12, g[0] - 1e-12, g[3]):
            p = pts_in[:60].copy()
            p[:, a] = val
            faces.append(p)
    pts = np.vstack([pts_in] + faces + [rng.uniform(-0.05, 0.05, (200, 3))])
    rec = {"pts": pts, "gx": grid["x"], "gy": grid["y"], "gz": grid["z"]}
    for fb in "EB":
        for cc in "xyz":
            rec[f"val_{fb}{cc}"] = vals[fb][cc]
    for backend in ("fast", "scipy"):
        fE = F.from_arrays(grid, vals["E"], scaling=1.25, interpolator_backend=backend)
        fB = F.from_arrays(grid, vals["B"], interpolator_backend=backend)
        rec[f"E3d_{backend}"] = fE(pts)
        rec[f"B3d_{backend}"] = fB(pts)
    # 1-D (z only) and 2-D (x, y) fields
    g1 = {"x": np.array([0.0]), "y": np.array([0.0]), "z": grid["z"]}
    v1 = {cc: vals["B"][cc][4, 5, :] for cc in "xyz"}
    rec["val1d"] = np.stack([v1[cc] for cc in "xyz"])
    # _get_field_1d hands the interpolator a flat (M,) array (field.py:466-470);
    # CoordinateMapper/NumbaInterpolator read that as ONE point of dimension M
    # and raise, so only the 'scipy' backend evaluates 1-D fields.
    try:
        F.from_arrays(g1, v1, scaling=0.5, interpolator_backend="fast")(pts)
        rec["fast1d_error"] = np.array("none")
    except Exception as exc:
        rec["fast1d_error"] = np.array(type(exc).__name__)
    rec["B1d"] = F.from_arrays(g1, v1, scaling=0.5, interpolator_backend="scipy")(pts)
    g2 = {"x": grid["x"], "y": grid["y"], "z": np.array([0.0])}
    v2 = {cc: vals["E"][cc][:, :, 7] for cc in "xyz"}
    rec["val2d"] = np.stack([v2[cc] for cc in "xyz"])
    rec["E2d"] = F.from_arrays(g2, v2)(pts)
    c0 = F(dim=0, field={"x": 0.3, "y": -1.1, "z": 2.5}, scaling=3.0)
    rec["const0d"] = c0(pts)
    fB = F.from_arrays(grid, vals["B"])
    rec["scaled"] = (2.5 * fB)(pts)
    rec["offset"] = (fB + 0.125)(pts)
    rec["composite"] = (fB - c0)(pts)
    rec["zero3"] = F.zero(3)(pts)
    np.savez(os.path.join(OUT, "field_call.npz"), **rec)

    # ---- track_batch: uniform fields, all algorithms ------------------------
    M = 64
    cases = {}
    for name, sp, beta_max in (("p", "proton", 0.08), ("e", "electron", 0.95)):
        d = rng.standard_normal((M, 3))
        d /= np.linalg.norm(d, axis=1)[:, None]
        v0 = d * (rng.uniform(0.01, beta_max, M) * c)[:, None]
        r0 = rng.uniform(-5e-3, 5e-3, (M, 3))
        cases[name] = (sp, r0, v0)
    Ec = {"x": 1.0e5, "y": -2.0e5, "z": 5.0e4}
    Bc = {"x": 0.1, "y": 0.2, "z": 1.0}
    rec = {}
    for name, (sp, r0, v0) in cases.items():
        rec[f"r0_{name}"], rec[f"v0_{name}"] = r0, v0
        rec[f"qm_{name}"] = S(sp).q_over_m
        dt = 1e-11 if name == "p" else 2e-13
        rec[f"dt_{name}"] = dt
        for alg in algs:
            pu = P(S(sp), alg)
            eF = F(dim=0, field=dict(Ec), scaling=1.5)
            bF = F(dim=0, field=dict(Bc))
            for nsteps, recn in ((200, 50), (200, 30)):
                rh, vh, act = pu.track_batch(r0, v0, eF, bF, nsteps, dt, recn)
                tag = f"{name}_{alg}_{nsteps}_{recn}"
                rec[f"r_{tag}"], rec[f"v_{tag}"], rec[f"a_{tag}"] = rh, vh, act
    rec["Ec"] = np.array([Ec[k] for k in "xyz"])
    rec["Bc"] = np.array([Bc[k] for k in "xyz"])
    rec["E_scaling"] = 1.5
    np.savez(os.path.join(OUT, "track_batch_uniform.npz"), **rec)

    # ---- track_batch: 16^3 map, 'fast' backend -------------------------------
    grid, vals = c2_map(16, 16, 16, 1e-3, rng)
    rec = {"gx": grid["x"], "gy": grid["y"], "gz": grid["z"]}
    for fb in "EB":
        for cc in "xyz":
            rec[f"val_{fb}{cc}"] = vals[fb][cc]
    M = 64
    r0 = np.column_stack([rng.uniform(-6e-3, 6e-3, M), rng.uniform(-6e-3, 6e-3, M), rng.uniform(1e-3, 14e-3, M)])
    d = rng.standard_normal((M, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    v0p = d * 1.0e6
    v0e = d * (0.6 * c)
    rec["r0"], rec["v0p"], rec["v0e"] = r0, v0p, v0e
    eF = F.from_arrays(grid, vals["E"], scaling=1.0)
    bF = F.from_arrays(grid, vals["B"], scaling=2.0)
    for alg in algs:
        for sp, v0, dt, tag in (("proton", v0p, 1e-10, "p"), ("electron", v0e, 2e-13, "e")):
            pu = P(S(sp), alg)
            rh, vh, act = pu.track_batch(r0, v0, eF, bF, 120, dt, 40)
            rec[f"r_{tag}_{alg}"], rec[f"v_{tag}_{alg}"] = rh, vh
            rec[f"qm_{tag}"], rec[f"dt_{tag}"] = S(sp).q_over_m, dt
    np.savez(os.path.join(OUT, "track_batch_map.npz"), **rec)

    # ---- single-particle track / push, push_batch ---------------------------
    rec = {}
    sp, r0, v0 = cases["p"]
    eF = F(dim=0, field=dict(Ec), scaling=1.5)
    bF = F(dim=0, field=dict(Bc))
    for alg in algs:
        pu = P(S(sp), alg)
        rh, vh = pu.track(r0[3], v0[3], eF, bF, 50, 1e-11, 7)
        rec[f"trk_r_{alg}"], rec[f"trk_v_{alg}"] = rh, vh
        rn, vn = pu.push(r0[5], v0[5], eF, bF, 1e-11)
        rec[f"push_r_{alg}"], rec[f"push_v_{alg}"] = rn, vn
        rn, vn = pu.push_batch(r0, v0, eF, bF, 1e-11)
        rec[f"pb_r_{alg}"], rec[f"pb_v_{alg}"] = rn, vn
    rec["r0"], rec["v0"], rec["qm"] = r0, v0, S(sp).q_over_m
    np.savez(os.path.join(OUT, "track_single.npz"), **rec)

    # ---- track_batch with particle loss (mock electrode wall) ---------------
    rec = {}
    M = 32
    r0 = rng.uniform(-2e-3, 2e-3, (M, 3))
    v0 = rng.standard_normal((M, 3)) * 2e5
    v0[:, 0] = -np.abs(v0[:, 0]) - 1e5          # most move away from the wall
    v0[-5:, 0] = 2.0e6                           # the last five hit it
    wall = 2.5e-3
    eF = F(dim=0, field={"x": 0.0, "y": 0.0, "z": 0.0})
    bF = F(dim=0, field={"x": 0.0, "y": 0.0, "z": 0.2})
    for alg in ("boris", "rk4", "vay_rel"):
        pu = P(S("proton"), alg, electrode_assembly=_WallAssembly(wall))
        rh, vh, act = pu.track_batch(r0, v0, eF, bF, 60, 1e-10, 10)
        rec[f"r_{alg}"], rec[f"v_{alg}"], rec[f"a_{alg}"] = rh, vh, act
    # A lost particle that is not among the last ones makes the no-numba Boris
    # finalisation index past the active-subset field array (pusher.py:1000-1005).
    v1 = v0.copy()
    v1[3, 0] = 2.0e6
    pu = P(S("proton"), "boris", electrode_assembly=_WallAssembly(wall))
    try:
        pu.track_batch(r0, v1, eF, bF, 60, 1e-10, 10)
        rec["middle_loss_error"] = np.array("none")
    except Exception as exc:  # recorded as a fact of the reference
        rec["middle_loss_error"] = np.array(type(exc).__name__)
    rec["r0"], rec["v0"], rec["v0_middle"], rec["wall"] = r0, v0, v1, wall
    rec["qm"] = S("proton").q_over_m
    np.savez(os.path.join(OUT, "track_batch_loss.npz"), **rec)
    print("golden fixtures written to", OUT)


if __name__ == "__main__":
    main()
