"""Beam moments of ParticleDistribution, restated in NumPy FP64 (test infrastructure only).

Follows src/PyPATools/particles.py:
  calculate_centroid / calculate_mean_pr / calculate_mean_energy_mev /
  calculate_stdevs (:693-723), get_beam_edges (:725-759), get_emittances
  (:761-788), get_twiss_parameters (:790-843), ekin_mev (:507-525), xp_vec
  (:463-467), mean_momentum_betagamma (:599-605), and the momentum <-> velocity
  conversions (:72-117, :393-457).

The moment functions take the velocity array explicitly, so the same formulas
serve both the reference's per-component conversion (``reference_v_from_p``)
and the physical one (``physical_v_from_p``) the device engine uses.
"""
import numpy as np

from .constants import CLIGHT

EPSILON = 1e-10   # global_variables.py:33


def reference_v_from_p(p):
    """particles.py:72-77, 393-402 (RELATIVISTIC): v_i = c p_i / sqrt(p_i^2 + 1), component by component."""
    p = np.asarray(p, dtype=np.float64)
    return CLIGHT * p / np.sqrt(p ** 2.0 + 1.0)


def reference_p_from_v(v):
    """particles.py:95-104, 428-434: p_i = v_i / sqrt(c^2 - v_i^2), component by component."""
    v = np.asarray(v, dtype=np.float64)
    return v / np.sqrt(CLIGHT ** 2 - v ** 2)


def physical_v_from_p(p):
    """v = c p / sqrt(1 + |p|^2) for p = beta gamma (the vector form)."""
    p = np.asarray(p, dtype=np.float64)
    return CLIGHT * p / np.sqrt(1.0 + np.sum(p * p, axis=1))[:, None]


def physical_p_from_v(v):
    """p = gamma v / c with gamma from |v| (the vector form)."""
    v = np.asarray(v, dtype=np.float64)
    b2 = np.sum(v * v, axis=1) / CLIGHT ** 2
    return v / CLIGHT / np.sqrt(1.0 - b2)[:, None]


def angles(v):
    """particles.py:463-467: v / v_z with |v_z| < 1e-10 replaced by EPSILON."""
    vz = np.where(np.abs(v[:, 2]) < 1e-10, EPSILON, v[:, 2])
    return v / vz[:, None]


def ekin_mev(p, mass_mev, z_energy=False):
    """particles.py:507-525 (RELATIVISTIC): sqrt((pr m)^2 + m^2) - m, pr = |p| or p_z."""
    pr = p[:, 2] if z_energy else np.linalg.norm(p, axis=1)
    return np.sqrt((pr * mass_mev) ** 2 + mass_mev ** 2) - mass_mev


def moments(x, p, v, mass_mev, z_energy=False):
    """The collective quantities of recalculate_all (particles.py:685-723) for arrays x, p, v (N, 3)."""
    xp = angles(v)
    e = ekin_mev(p, mass_mev, z_energy)
    c = np.mean(x, axis=0)
    xpm = np.mean(xp, axis=0)
    return {
        "n": len(x), "centroid": c, "mean_momentum": np.mean(p, axis=0),
        "mean_betagamma": np.mean(p[:, 2]) if z_energy else np.mean(np.linalg.norm(p, axis=1)),
        "ekin_mean": np.mean(e), "ekin_std": np.std(e),
        "x_std": np.std(x, axis=0), "xp_std": np.std(xp, axis=0), "xp_mean": xpm,
        "xxp": np.mean((x[:, 0] - c[0]) * (xp[:, 0] - xpm[0])),
        "yyp": np.mean((x[:, 1] - c[1]) * (xp[:, 1] - xpm[1])),
        "xyp": np.mean((x[:, 0] - c[0]) * (xp[:, 1] - xpm[1])),
        "yxp": np.mean((x[:, 1] - c[1]) * (xp[:, 0] - xpm[0])),
        "min": np.min(x, axis=0), "max": np.max(x, axis=0),
    }


def beam_edges(m, mode="1rms"):
    """particles.py:725-759."""
    if mode == "1rms":
        lo, hi = m["centroid"] - m["x_std"], m["centroid"] + m["x_std"]
    elif mode == "2rms":
        lo, hi = m["centroid"] - 2.0 * m["x_std"], m["centroid"] + 2.0 * m["x_std"]
    elif mode == "full":
        lo, hi = m["min"], m["max"]
    else:
        raise ValueError(f"Unknown mode: {mode}. Use '1rms', '2rms', or 'full'")
    return np.column_stack((lo, hi)).flatten()


def emittances(m, normalized=True):
    """particles.py:761-788: [e_xx', e_yy', e_xy', e_yx'], times mean beta*gamma when normalized."""
    sx, sxp = m["x_std"], m["xp_std"]
    e = np.array([np.sqrt((sx[0] * sxp[0]) ** 2 - m["xxp"] ** 2), np.sqrt((sx[1] * sxp[1]) ** 2 - m["yyp"] ** 2),
                  np.sqrt((sx[0] * sxp[1]) ** 2 - m["xyp"] ** 2), np.sqrt((sx[1] * sxp[0]) ** 2 - m["yxp"] ** 2)])
    return e * m["mean_betagamma"] if normalized else e


def twiss(m):
    """particles.py:790-818: [alpha_x, beta_x, gamma_x, alpha_y, beta_y, gamma_y]."""
    ex, ey, _, _ = emittances(m, normalized=False)
    bx, gx = m["x_std"][0] ** 2 / ex, m["xp_std"][0] ** 2 / ex
    ax = np.sqrt(bx * gx - 1.0) if m["xxp"] < 0 else -np.sqrt(bx * gx - 1.0)
    by, gy = m["x_std"][1] ** 2 / ey, m["xp_std"][1] ** 2 / ey
    ay = np.sqrt(by * gy - 1.0) if m["yyp"] < 0 else -np.sqrt(by * gy - 1.0)
    return np.array([ax, bx, gx, ay, by, gy])


def ellipse_fractions(x, v, tw, ex, ey):
    """particles.py:821-835: percent of particles with gamma x^2 + 2 alpha x x' + beta x'^2 < 4 eps."""
    xp = angles(v)
    fx = tw[2] * x[:, 0] ** 2 + 2.0 * tw[0] * x[:, 0] * xp[:, 0] + tw[1] * xp[:, 0] ** 2
    fy = tw[5] * x[:, 1] ** 2 + 2.0 * tw[3] * x[:, 1] * xp[:, 1] + tw[4] * xp[:, 1] ** 2
    return 100.0 * np.sum(fx < 4.0 * ex) / len(x), 100.0 * np.sum(fy < 4.0 * ey) / len(x)
