"""Synthetic workloads (seeded) for the benchmark and the parity tests.

C1  uniform B = 1 T along z, E = 0; 1e4 protons at 1 MeV, isotropic directions,
    positions uniform in a 1 cm cube; dt = 1e-11 s, 1000 steps; 'boris'.
C2  128^3 map at 1 mm spacing, x, y in [-63.5, 63.5] mm, z in [0, 127] mm,
    periodic in z with L = 128 mm: Bz = 0.5 T (1 + 0.3 cos(2 pi z / L)),
    paraxial Bx,y = -(x,y)/2 dBz/dz, E = 2 MV/m^2 (x, -y, 0); 1e6 protons in a
    Gaussian bunch (sigma = 2 mm in x, y, z) at the map centre, |p| =
    p(2.5 MeV) (1 + 1e-3 N(0,1)) along +z; dt = 5e-11 s, 1e4 steps.
C3  96 x 96 x 256 TM010-like map (E0 = 10 MV/m, f = 100 MHz, cos/sin(w t + phi),
    phi from the bunch seed), x, y in [-9.5, 9.5] mm (0.2 mm), z in [0, 318.75] mm
    (1.25 mm: longer than c N dt = 300 mm, so no electron leaves the map along z
    in the 5000 steps; they reach about 1.4 MeV and z = 260 mm); 4e6 electrons at
    100 keV along +z in a uniform disk of radius 3 mm at z = 10 mm; dt = 2e-13 s,
    5000 steps; 'rk4_rel'.
All particle inputs are float64 (M, 3) arrays in SI units, drawn with
numpy.random.default_rng(seed).
"""
import numpy as np

from .constants import CLIGHT
from .species import IonSpecies


def velocity_from_kinetic(ke_mev, species, n=None):
    """|v| of a particle of kinetic energy ke_mev [MeV]."""
    m = IonSpecies(species).mass_mev
    g = 1.0 + ke_mev / m
    return np.sqrt(1.0 - 1.0 / g ** 2) * CLIGHT


def c1(M=10_000, seed=1):
    rng = np.random.default_rng(seed)
    speed = velocity_from_kinetic(1.0, "proton")
    d = rng.standard_normal((M, 3))
    v0 = d / np.linalg.norm(d, axis=1)[:, None] * speed
    r0 = rng.uniform(-5e-3, 5e-3, (M, 3))
    return {"r0": r0, "v0": v0, "E": (0.0, 0.0, 0.0), "B": (0.0, 0.0, 1.0), "dt": 1e-11, "nsteps": 1000,
            "species": "proton", "algorithm": "boris"}


def c2_map(n=128, dx=1e-3):
    x = (np.arange(n) - (n - 1) / 2.0) * dx
    z = np.arange(n) * dx
    L = n * dx
    X, Y, Z = np.meshgrid(x, x, z, indexing="ij")
    k = 2 * np.pi / L
    bz = 0.5 * (1.0 + 0.3 * np.cos(k * Z))
    dbz = -0.5 * 0.3 * k * np.sin(k * Z)
    E = {"x": 2e6 * X, "y": -2e6 * Y, "z": np.zeros_like(X)}
    B = {"x": -0.5 * X * dbz, "y": -0.5 * Y * dbz, "z": bz}
    return {"x": x, "y": x, "z": z}, E, B


def c2_bunch(M=1_000_000, seed=2, n=128, dx=1e-3):
    rng = np.random.default_rng(seed)
    zc = (n - 1) / 2.0 * dx
    r0 = np.column_stack([rng.normal(0.0, 2e-3, M), rng.normal(0.0, 2e-3, M), rng.normal(zc, 2e-3, M)])
    m = IonSpecies("proton").mass_mev
    g = 1.0 + 2.5 / m
    u0 = np.sqrt(g * g - 1.0) * CLIGHT                       # gamma * beta * c at 2.5 MeV
    u = u0 * (1.0 + 1e-3 * rng.standard_normal(M))
    vz = u / np.sqrt(1.0 + (u / CLIGHT) ** 2)
    v0 = np.column_stack([np.zeros(M), np.zeros(M), vz])
    return r0, v0


def c2(M=1_000_000, seed=2, n=128):
    grid, E, B = c2_map(n)
    r0, v0 = c2_bunch(M, seed, n)
    return {"grid": grid, "E": E, "B": B, "r0": r0, "v0": v0, "dt": 5e-11, "nsteps": 10_000,
            "species": "proton", "periodic_z": True}


def c3_map(nx=96, nz=256, dxy=0.2e-3, dz=1.25e-3, E0=1e7, f=1e8):
    from scipy.special import j0, j1
    x = (np.arange(nx) - (nx - 1) / 2.0) * dxy
    z = np.arange(nz) * dz
    X, Y, Z = np.meshgrid(x, x, z, indexing="ij")
    k = 2 * np.pi * f / CLIGHT
    r = np.hypot(X, Y)
    ez = E0 * j0(k * r)
    bth = E0 / CLIGHT * j1(k * r)
    with np.errstate(invalid="ignore", divide="ignore"):
        cx = np.where(r > 0, X / np.where(r > 0, r, 1.0), 0.0)
        cy = np.where(r > 0, Y / np.where(r > 0, r, 1.0), 0.0)
    E = {"x": np.zeros_like(X), "y": np.zeros_like(X), "z": ez}
    B = {"x": -bth * cy, "y": bth * cx, "z": np.zeros_like(X)}
    return {"x": x, "y": x, "z": z}, E, B, 2 * np.pi * f


def c3_bunch(M=4_000_000, seed=3):
    rng = np.random.default_rng(seed)
    rad = 3e-3 * np.sqrt(rng.uniform(0, 1, M))
    ang = rng.uniform(0, 2 * np.pi, M)
    r0 = np.column_stack([rad * np.cos(ang), rad * np.sin(ang), np.full(M, 10e-3)])
    speed = velocity_from_kinetic(0.1, "electron")
    v0 = np.column_stack([np.zeros(M), np.zeros(M), np.full(M, speed)])
    phase = float(rng.uniform(0, 2 * np.pi))
    return r0, v0, phase
