"""Field-map file loaders (src/PyPATools/field_src/field_loaders.py).

Each loader parses one file format into the dictionary the reference's loaders
return: ``grid`` (1-D axes in metres), ``values`` (component arrays in T or V/m,
indexed [ix, iy, iz] on the axes present), ``dim``, ``field_type`` and
``metadata``.  ``Field.from_file`` turns the dictionary into a Field whose map is
uploaded to the device on first use.

Parsing is vectorised (one ``numpy`` conversion of the whole data block) rather
than line by line; blank lines in a data block are ignored.  The unit table,
header grammar, axis ordering and error types follow the reference.
"""
import os
import re

import numpy as np

# field_loaders.py:18-36: factors to SI (metres, tesla, V/m)
UNIT_CONVERSIONS = {
    "length": {"mm": 0.001, "cm": 0.01, "m": 1.0},
    "magnetic": {"gauss": 1e-4, "kgauss": 0.1, "tesla": 1.0, "t": 1.0},
    "electric": {"v/mm": 1000.0, "v/cm": 100.0, "v/m": 1.0},
}


def _numbers(text):
    """All whitespace-separated numbers of a text block as one float64 array."""
    return np.array(text.split(), dtype=np.float64)


def _cyl_to_cart_components(br, bt, theta_deg):
    """(B_r, B_theta) at angle theta [deg] -> (B_x, B_y)  (field_loaders.py:53-62)."""
    th = np.deg2rad(theta_deg)
    c, s = np.cos(th), np.sin(th)
    return br * c - bt * s, br * s + bt * c


def load_opera_table(filename, extents=None, extents_dims=None):
    """OPERA .table map (field_loaders.py:65-256).

    Header: the point counts, then ``<col> <label> [<unit>]`` lines up to a line
    ``0``; then one row per grid point.  Grid axes come from the distinct values of
    the X/Y/Z columns (or from ``extents`` in cm when the file has none); field
    columns are reshaped to (nX, nY, nZ) in file order.
    """
    if not os.path.exists(filename):
        raise FileNotFoundError(f"File not found: {filename}")
    if extents is not None and extents_dims is None:
        raise ValueError("extents_dims must be specified when using extents")
    try:
        with open(filename, "r") as fh:
            counts = np.array([int(t) - 1 for t in fh.readline().split()])[:3]
            cols = {}
            n_len = n_b = n_e = 0
            last_col = -1
            while True:
                line = fh.readline()
                if line == "":
                    raise ValueError("header not terminated by a '0' line")
                line = line.strip()
                if line == "0":
                    break
                parts = line.split()
                last_col = int(parts[0]) - 1
                label = parts[1]
                m = re.search(r"\[(.*?)]", line)
                if not m:
                    raise ValueError(f"No unit found for column {label}")
                unit = m.group(1)
                cols[label] = {"column": last_col, "unit": unit}
                u = unit.lower()
                if u in UNIT_CONVERSIONS["length"]:
                    n_len += 1
                elif u in UNIT_CONVERSIONS["magnetic"]:
                    n_b += 1
                elif u in UNIT_CONVERSIONS["electric"]:
                    n_e += 1
            if extents is not None:
                for k, label in enumerate(extents_dims):
                    cols[label] = {"column": last_col + k + 1, "unit": "cm"}
                    n_len += 1
            if "X" in cols:
                n = {"X": counts[0] + 1, "Y": counts[1] + 1 if "Y" in cols else 1,
                     "Z": counts[2] + 1 if "Z" in cols else 1}
            elif "Y" in cols:
                n = {"X": 1, "Y": counts[0] + 1, "Z": counts[1] + 1 if "Z" in cols else 1}
            elif "Z" in cols:
                n = {"X": 1, "Y": 1, "Z": counts[0] + 1}
            else:
                raise ValueError("no X, Y or Z column")
            n = {k: int(v) for k, v in n.items()}
            dim = int(np.count_nonzero(counts > 0))
            if n_e > 0 and n_b > 0:
                raise ValueError("Mixed E and B fields not supported")
            field_type = "electric" if n_e > 0 else "magnetic"
            npts = n["X"] * n["Y"] * n["Z"]
            ncols = n_len + n_e + n_b
            body = _numbers(fh.read())
        if extents is not None:
            nfile = ncols - n_len                        # the file holds the field columns only
            vals = body.reshape(npts, nfile)
            gx = np.linspace(extents[0][0], extents[0][1], n["X"])
            gy = np.linspace(extents[1][0], extents[1][1], n["Y"])
            gz = np.linspace(extents[2][0], extents[2][1], n["Z"])
            X, Y, Z = np.meshgrid(gx, gy, gz, indexing="ij")
            raw = np.empty((npts, ncols))
            raw[:, :nfile] = vals
            raw[:, -3], raw[:, -2], raw[:, -1] = X.ravel(), Y.ravel(), Z.ravel()
        else:
            if body.size != npts * ncols:
                raise ValueError(f"expected {npts} rows of {ncols} columns, found {body.size} numbers")
            raw = body.reshape(npts, ncols)
    except Exception as e:
        raise RuntimeError(f"Error reading OPERA table file: {e}")
    raw = raw.T
    grid = {}
    for key in ("X", "Y", "Z"):
        if key in cols and cols[key]["unit"].lower() in UNIT_CONVERSIONS["length"]:
            grid[key.lower()] = np.unique(raw[cols[key]["column"]]) * UNIT_CONVERSIONS["length"][
                cols[key]["unit"].lower()]
    kind = "magnetic" if field_type == "magnetic" else "electric"
    labels = ("BX", "BY", "BZ") if field_type == "magnetic" else ("EX", "EY", "EZ")
    values = {}
    for label in labels:
        comp = label[-1].lower()
        if label in cols:
            scale = UNIT_CONVERSIONS[kind].get(cols[label]["unit"].lower(), 1.0)
            values[comp] = raw[cols[label]["column"]].reshape(n["X"], n["Y"], n["Z"]) * scale
        else:
            values[comp] = np.zeros((n["X"], n["Y"], n["Z"]))
    return {"grid": grid, "values": values, "dim": dim, "field_type": field_type,
            "metadata": {"n_points": n, "units_original": {k: v["unit"] for k, v in cols.items()},
                         "filename": os.path.basename(filename)}}


# COMSOL column names -> canonical labels (field_loaders.py:304-312)
_COMSOL_LABELS = {
    "x": "X", "y": "Y", "z": "Z", "r": "R", "phi": "PHI", "theta": "THETA",
    "mf.Bx": "BX", "mf.By": "BY", "mf.Bz": "BZ", "Bx": "BX", "By": "BY", "Bz": "BZ",
    "mf.Br": "BR", "mf.Bphi": "BPHI", "mf.Btheta": "BTHETA",
    "es.Ex": "EX", "es.Ey": "EY", "es.Ez": "EZ", "es.Er": "ER", "es.Ephi": "EPHI", "es.Etheta": "ETHETA",
}
_SPATIAL = ("X", "Y", "Z", "R", "PHI", "THETA")
_B_LABELS = ("BX", "BY", "BZ", "BR", "BPHI", "BTHETA")
_E_LABELS = ("EX", "EY", "EZ", "ER", "EPHI", "ETHETA")
_COMSOL_OUT = {"BX": "x", "BY": "y", "BZ": "z", "BR": "r", "BPHI": "phi", "BTHETA": "theta",
               "EX": "x", "EY": "y", "EZ": "z", "ER": "r", "EPHI": "phi", "ETHETA": "theta"}


def load_comsol(filename):
    """COMSOL text export (field_loaders.py:259-486).

    A 9-line ``%`` header (dimension on line 4, node count on line 5, length unit on
    line 8, column names on line 9), then one row per node in any order.  Rows are
    ordered by the spatial columns (first column slowest) and reshaped onto the grid
    of their distinct values.
    """
    if not os.path.exists(filename):
        raise FileNotFoundError(f"File not found: {filename}")
    try:
        with open(filename, "r") as fh:
            header = [fh.readline().strip() for _ in range(9)]
            body = fh.read()
        array_len = int(header[4].split()[2])
        l_unit = header[7].split()[3]
        names = header[8]
        b_unit = e_unit = None
        if "(T)" in names:
            b_unit, field_type = "T", "magnetic"
        elif "(V/cm)" in names:
            e_unit, field_type = "V/cm", "electric"
        elif "(V/m)" in names:
            e_unit, field_type = "V/m", "electric"
        elif "mf.B" in names:
            b_unit, field_type = "T", "magnetic"
        elif "es.E" in names:
            e_unit, field_type = "V/m", "electric"
        else:
            raise ValueError("Cannot determine field type from header")
        toks = names.split()
        toks = toks[:4] + re.split(r"\([a-zA-Z/]+\)", "".join(toks[4:]))
        cols, spatial, fields = {}, [], []
        j = 0
        for tok in toks:
            if tok in ("%", "(T)", "(V/cm)", "(V/m)", "", " "):
                continue
            name = re.sub(r"side|mir|sec|cpl|\d+", "", tok)
            if name not in _COMSOL_LABELS:
                continue
            lab = _COMSOL_LABELS[name]
            cols[lab] = {"column": j}
            if lab in _SPATIAL:
                cols[lab]["unit"] = l_unit
                spatial.append(lab)
            elif lab in _B_LABELS and b_unit:
                cols[lab]["unit"] = b_unit
                fields.append(lab)
            elif lab in _E_LABELS and e_unit:
                cols[lab]["unit"] = e_unit
                fields.append(lab)
            j += 1
        if not spatial:
            raise ValueError("No spatial coordinates found in file")
        if not fields:
            raise ValueError("No field components found in file")
        dim = len(spatial)
        first = next((ln for ln in body.splitlines() if ln.strip()), "")
        ncol = len(first.split())
        rows = _numbers(body)[: array_len * ncol].reshape(array_len, ncol)
    except Exception as e:
        raise RuntimeError(f"Error reading COMSOL file: {e}")
    sidx = [cols[c]["column"] for c in spatial]
    order = np.argsort(rows[:, sidx[0]], kind="stable") if dim == 1 else \
        np.lexsort([rows[:, sidx[k]] for k in range(dim - 1, -1, -1)])
    rows = rows[order]
    l_scale = UNIT_CONVERSIONS["length"].get(l_unit.lower(), 1.0)
    n, grid = {}, {}
    for c in spatial:
        u = np.unique(rows[:, cols[c]["column"]])
        n[c] = len(u)
        grid[c.lower()] = u * l_scale
    if int(np.prod(list(n.values()))) != array_len:
        raise ValueError(f"Grid size mismatch: {int(np.prod(list(n.values())))} expected vs {array_len} in file")
    shape = [n[c] for c in spatial]
    kind = "magnetic" if field_type == "magnetic" else "electric"
    values = {}
    for lab in fields:
        scale = UNIT_CONVERSIONS[kind].get(cols[lab]["unit"].lower(), 1.0)
        values[_COMSOL_OUT.get(lab, lab.lower())] = rows[:, cols[lab]["column"]].reshape(shape) * scale
    for comp in ("x", "y", "z"):
        values.setdefault(comp, np.zeros(shape))
    for ax in ("x", "y", "z"):
        grid.setdefault(ax, np.array([0.0]))
    return {"grid": grid, "values": values, "dim": dim, "field_type": field_type,
            "metadata": {"n_points": n, "spatial_dims": spatial, "field_dims": fields,
                         "units_original": {k: v["unit"] for k, v in cols.items() if "unit" in v},
                         "filename": os.path.basename(filename)}}


def _rgi2(r_axis, th_axis, table, r_pts, th_pts):
    """Bilinear interpolation on a regular (r, theta) grid, 0 outside (scipy RGI 'linear')."""
    from scipy.interpolate import RegularGridInterpolator
    f = RegularGridInterpolator((r_axis, th_axis), table, method="linear", bounds_error=False, fill_value=0.0)
    return f(np.column_stack([r_pts.ravel(), th_pts.ravel()])).reshape(r_pts.shape)


def _polar_of(X, Y):
    """Radius and angle in [0, 360) degrees of Cartesian points."""
    th = np.rad2deg(np.arctan2(Y, X))
    return np.hypot(X, Y), np.where(th < 0, th + 360.0, th)


def load_opal_midplane(filename, r_cutoff=None, scaling=1.0, cartesian_grid=True, nx=None, ny=None):
    """OPAL CARBONCYCL midplane Bz map, .dat (field_loaders.py:489-670).

    Six header lines (r0 [mm], dr [mm], theta0 [deg], dtheta [deg], n_theta, n_r),
    then Bz in kGauss with theta fastest.  Returned on the native (r, theta) grid
    or resampled bilinearly onto an nx x ny Cartesian midplane grid.
    """
    if not os.path.exists(filename):
        raise FileNotFoundError(f"File not found: {filename}")
    if os.path.splitext(filename)[1] != ".dat":
        raise ValueError("File extension must be '.dat'")
    try:
        with open(filename, "r") as fh:
            sr = float(fh.readline().strip())
            dr = float(fh.readline().strip())
            sth = float(fh.readline().strip())
            dth = float(fh.readline().strip())
            nth = int(fh.readline().strip())
            nr = int(fh.readline().strip())
            body = fh.read()
    except Exception as e:
        raise RuntimeError(f"Error reading OPAL midplane file: {e}")
    bz = scaling * 0.1 * _numbers(body)
    if bz.size != nth * nr:
        print(f"Warning: Found {bz.size} Bz values, expected {nth * nr}")
    r_axis = np.linspace(sr, sr + (nr - 1) * dr, nr)          # mm
    th_axis = np.linspace(sth, sth + (nth - 1) * dth, nth)    # deg
    bz = bz.reshape(nr, nth)
    if r_cutoff is not None:
        bz[r_axis > r_cutoff * 1000.0, :] = 0.0
    meta = {"r_range": (sr, sr + (nr - 1) * dr), "theta_range": (sth, sth + (nth - 1) * dth),
            "n_points": {"r": nr, "theta": nth}, "r_cutoff": r_cutoff, "scaling": scaling,
            "filename": os.path.basename(filename)}
    if not cartesian_grid:
        return {"grid": {"r": r_axis * 0.001, "theta": th_axis, "z": np.array([0.0])},
                "values": {"r": np.zeros_like(bz), "theta": np.zeros_like(bz), "z": bz},
                "dim": 2, "field_type": "magnetic", "coordinate_system": "cylindrical", "metadata": meta}
    r_max = r_axis[-1] * 0.001
    if r_cutoff is not None:
        r_max = min(r_max, r_cutoff)
    if nx is None:
        nx = int(2 * r_max / (dr * 0.001)) + 1
    if ny is None:
        ny = nx
    xs = np.linspace(-r_max, r_max, nx)
    ys = np.linspace(-r_max, r_max, ny)
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    rr, th = _polar_of(X, Y)
    bzc = _rgi2(r_axis, th_axis, bz, rr * 1000.0, th)
    if r_cutoff is not None:
        bzc[rr > r_cutoff] = 0.0
    meta["n_cartesian"] = {"x": nx, "y": ny}
    return {"grid": {"x": xs, "y": ys, "z": np.array([0.0])},
            "values": {"x": np.zeros_like(bzc), "y": np.zeros_like(bzc), "z": bzc},
            "dim": 2, "field_type": "magnetic", "coordinate_system": "cartesian", "metadata": meta}


def load_aima_agora(path, mirror=False, interp_resolution=1000):
    """AIMA AGORA cyclotron map: one .map file per (z, component) (field_loaders.py:673-888).

    File names ``...Z<z mm>-<br|bf|bz>.map``; each file has a comment line,
    ``n_theta n_r``, ``r0 dr`` [cm], ``theta0 dtheta`` [deg], then values with
    theta fastest (the last line is not data).  Every z slice is resampled onto an
    interp_resolution^2 Cartesian grid and (B_r, B_theta) rotated to (B_x, B_y);
    ``mirror`` adds the z < 0 half with B_x, B_y odd and B_z even.
    """
    if os.path.isfile(path):
        directory = os.path.dirname(path)
    elif os.path.isdir(path):
        directory = path
    else:
        raise FileNotFoundError(f"Path not found: {path}")
    try:
        names = [e.name for e in os.scandir(directory) if e.is_file() and e.name.endswith(".map")]
    except Exception as e:
        raise RuntimeError(f"Error scanning directory: {e}")
    if not names:
        raise FileNotFoundError(f"No .map files found in {directory}")
    entries = []
    for name in names:
        try:
            zs, rest = name.split("Z")[1].split("-")
            comp = rest.split(".")[0].lower()
            if comp not in ("br", "bf", "bz"):
                print(f"Warning: Skipping file with unknown component: {name}")
                continue
            entries.append((os.path.join(directory, name), comp, float(zs)))
        except Exception as e:
            print(f"Warning: Could not parse filename {name}: {e}")
    if not entries:
        raise ValueError("No valid .map files found")
    zsets = [{z for _, c, z in entries if c == comp} for comp in ("br", "bf", "bz")]
    if not (zsets[0] == zsets[1] == zsets[2]):
        raise ValueError("Not all components have the same z positions. Maybe a file is missing?")

    def header(fn):
        with open(fn, "r") as fh:
            fh.readline()
            nth_, nr_ = (int(v) for v in fh.readline().split())
            sr_, dr_ = (float(v) for v in fh.readline().split())
            sth_, dth_ = (float(v) for v in fh.readline().split())
        return nth_, nr_, sr_, dr_, sth_, dth_

    try:
        hdr = header(entries[0][0])
    except Exception as e:
        raise RuntimeError(f"Error reading grid parameters from {entries[0][0]}: {e}")
    for fn, _, _ in entries[1:]:
        if header(fn) != hdr:
            raise ValueError(f"Grid mismatch in file {fn}")
    nth, nr, sr, dr, sth, dth = hdr
    r_axis = np.linspace(sr, nr * dr, nr, endpoint=False)       # cm
    th_axis = np.linspace(sth, nth * dth, nth, endpoint=False)  # deg
    z_axis = np.sort(np.unique([z for _, _, z in entries])) / 10.0   # cm
    r_max = r_axis[-1]
    xs = np.linspace(-r_max, r_max, interp_resolution)
    ys = np.linspace(-r_max, r_max, interp_resolution)
    if mirror:
        zs_all = np.concatenate((-z_axis[:0:-1], z_axis))
        off = len(z_axis) - 1
    else:
        zs_all = z_axis
        off = 0
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    rr, th = _polar_of(X, Y)
    shape = (interp_resolution, interp_resolution, len(zs_all))
    bx, by, bz = np.zeros(shape), np.zeros(shape), np.zeros(shape)
    for j, zc in enumerate(z_axis):
        comp_tab = {}
        for comp in ("br", "bf", "bz"):
            match = [fn for fn, c, z in entries if c == comp and z == 10.0 * zc]
            if not match:
                raise ValueError(f"Missing {comp} component at z={10.0 * zc} mm")
            with open(match[0], "r") as fh:
                lines = fh.readlines()[4:-1]
            comp_tab[comp] = _numbers(" ".join(lines)).reshape(nr, nth)
        br = _rgi2(r_axis, th_axis, comp_tab["br"], rr, th)
        bf = _rgi2(r_axis, th_axis, comp_tab["bf"], rr, th)
        i = j + off
        bx[:, :, i], by[:, :, i] = _cyl_to_cart_components(br, bf, th)
        bz[:, :, i] = _rgi2(r_axis, th_axis, comp_tab["bz"], rr, th)
        if mirror and j != 0:
            k = off - j
            bx[:, :, k], by[:, :, k], bz[:, :, k] = -bx[:, :, i], -by[:, :, i], bz[:, :, i]
    return {"grid": {"x": xs * 0.01, "y": ys * 0.01, "z": zs_all * 0.01},
            "values": {"x": bx, "y": by, "z": bz}, "dim": 3, "field_type": "magnetic",
            "metadata": {"z_positions": z_axis, "n_files": len(names), "mirrored": mirror,
                         "interp_resolution": interp_resolution, "r_range": (sr, nr * dr),
                         "theta_range": (sth, nth * dth), "filename": os.path.basename(directory)}}


def load_h5part(filename):
    """H5Part (HDF5) map: the H field of Step#0/Block, stored [iz][iy][ix] (field_loaders.py:931-1049).

    Needs h5py, which is optional here; without it this raises ImportError.
    """
    if not os.path.exists(filename):
        raise FileNotFoundError(f"File not found: {filename}")
    try:
        import h5py
    except ImportError as e:
        raise ImportError("load_h5part needs h5py, which is not installed") from e
    try:
        with h5py.File(filename, "r") as h5:
            block = h5["Step#0"]["Block"]
            has_e = "Efield" in block
            if "Hfield" not in block:
                raise ValueError("No Hfield found in H5Part file")
            g = block["Hfield"]
            comps = [np.array(g[k]) for k in ("0", "1", "2")]
            spacing = np.asarray(g.attrs["__Spacing__"])
            origin = np.asarray(g.attrs["__Origin__"])
    except Exception as e:
        raise RuntimeError(f"Error reading H5Part file: {e}")
    nz, ny, nx = comps[0].shape
    axes = [origin[a] + np.arange(m) * spacing[a] for a, m in enumerate((nx, ny, nz))]
    vals = [np.transpose(c, (2, 1, 0)) for c in comps]
    return {"grid": dict(zip("xyz", axes)), "values": dict(zip("xyz", vals)), "dim": 3, "field_type": "magnetic",
            "metadata": {"spacing": spacing, "origin": origin, "shape": (nx, ny, nz), "has_efield": has_e,
                         "filename": os.path.basename(filename)}}


# Field.from_file dispatch by extension (field.py:305-318)
LOADERS = {".h5": load_h5part, ".h5part": load_h5part, ".table": load_opera_table, ".comsol": load_comsol,
           ".dat": load_opal_midplane, ".map": load_aima_agora}
