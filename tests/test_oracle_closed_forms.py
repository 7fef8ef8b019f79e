"""Pin the oracle with mathematics that does not restate its own formulas.

Closed forms (discrete Boris rotation, constant-E kicks, gyro-orbits, E x B
drift, 1-D RF momentum gain), convergence orders, library routines
(scipy RegularGridInterpolator / map_coordinates) and exactness of trilinear
interpolation on multilinear fields.  Each is chosen so that a dropped term, a
wrong sign or a swapped operand in the oracle fails it.
"""
import numpy as np
import pytest
from scipy.interpolate import RegularGridInterpolator

from oracle import kernels as K
from oracle import species as OS
from oracle.constants import CLIGHT
from oracle.field import OField, OPeriodicZ, ORF
from oracle.interp import interp_coordinate_mapper, interp_nd
from oracle.pusher import OraclePusher

QM_P = OS.q_over_m("proton")


def const(v):
    return OField.constant(*v)


def test_q_over_m_against_codata():
    # CODATA 2022 e/m_p = 9.578833143e7 C/kg; the reference uses a*AMU_MEV
    # (a = 1.00727647) which sits 1.4e-8 away from the proton mass.
    assert abs(QM_P / 9.578833143e7 - 1.0) < 2e-8
    # electron: e/m_e = 1.75882000838e11 C/kg, negative charge
    assert abs(OS.q_over_m("electron") / -1.75882000838e11 - 1.0) < 1e-10
    assert OS.q_over_m("H2_1+") == pytest.approx(QM_P * 1.00727647 / 2.01533, rel=1e-12)


def test_boris_pure_B_is_exact_rotation():
    """E = 0, uniform B: each step rotates v by theta = 2 atan(|q/m| B dt / 2)."""
    rng = np.random.default_rng(1)
    M, N, dt, Bz = 16, 400, 1e-11, 1.0
    v0 = rng.standard_normal((M, 3)) * 1e7
    r0 = rng.standard_normal((M, 3)) * 1e-3
    rh, vh, _ = OraclePusher(QM_P, "boris").track_batch(r0, v0, const((0, 0, 0)), const((0, 0, Bz)), N, dt, 1)
    theta = 2.0 * np.arctan(QM_P * Bz * dt / 2.0)
    # the half step back rotates by theta/2 with the same rule (t -> t/2)
    th_half = 2.0 * np.arctan(QM_P * Bz * dt / 4.0)
    w0 = (v0[:, 0] + 1j * v0[:, 1]) * np.exp(1j * th_half)   # q>0 rotates clockwise in +B
    n = np.arange(N + 1)[:, None]
    w = w0[None, :] * np.exp(-1j * theta * n)                 # v_{n-1/2}
    vh_c = vh[:-1, :, 0] + 1j * vh[:-1, :, 1]
    assert np.max(np.abs(vh_c - w[:-1])) / 1e7 < 1e-12
    assert np.allclose(vh[:, :, 2], v0[:, 2][None, :], rtol=0, atol=1e-6)
    # positions: r_n = r_0 + dt * sum_{k=1..n} v_{k-1/2}, v_{k-1/2} = w[k]
    zpos = (r0[:, 0] + 1j * r0[:, 1])[None, :] + dt * np.cumsum(np.vstack([np.zeros((1, M)), w[1:]]), axis=0)
    assert np.max(np.abs(rh[:, :, 0] + 1j * rh[:, :, 1] - zpos)) < 1e-15 + 1e-12 * np.max(np.abs(zpos))
    # the last record is resynchronised to integer time: rotated back by theta/2 from v_{N-1/2}
    wN = w[N] * np.exp(-1j * th_half)
    assert np.max(np.abs(vh[-1, :, 0] + 1j * vh[-1, :, 1] - wN)) / 1e7 < 1e-12
    # the exact rotation angle differs from omega*dt at third order
    assert abs(theta - QM_P * Bz * dt) < (QM_P * Bz * dt) ** 3


def test_boris_pure_E_constant_kick():
    M, N, dt = 8, 100, 1e-11
    rng = np.random.default_rng(2)
    E = np.array([3e5, -1e5, 2e5])
    v0 = rng.standard_normal((M, 3)) * 1e5
    r0 = np.zeros((M, 3))
    rh, vh, _ = OraclePusher(QM_P, "boris").track_batch(r0, v0, const(E), const((0, 0, 0)), N, dt, N)
    a = QM_P * E
    # v_{-1/2} = v0 - a dt/2 ; after N steps + half step forward: v0 + a N dt
    assert np.allclose(vh[-1], v0 + a * N * dt, rtol=1e-13, atol=1e-6)
    # r_N = r0 + sum_k v_{k-1/2} dt = r0 + N dt v0 + a dt^2 N^2 / 2
    assert np.allclose(rh[-1], v0 * N * dt + 0.5 * a * (N * dt) ** 2, rtol=1e-12, atol=1e-18)


def test_leapfrog_and_rk4_uniform_E():
    dt, N = 1e-10, 50
    E = np.array([1e4, 0.0, 0.0])
    a = QM_P * E
    r0 = np.zeros((1, 3))
    v0 = np.array([[0.0, 1e4, 0.0]])
    # RK4 integrates constant acceleration exactly
    rh, vh, _ = OraclePusher(QM_P, "rk4").track_batch(r0, v0, const(E), const((0, 0, 0)), N, dt, N)
    T = N * dt
    assert np.allclose(rh[-1], v0 * T + 0.5 * a * T ** 2, rtol=1e-12, atol=1e-20)
    # Euler-Cromer: r_N = v0 T + a dt^2 N(N+1)/2
    rh, vh, _ = OraclePusher(QM_P, "leapfrog").track_batch(r0, v0, const(E), const((0, 0, 0)), N, dt, N)
    assert np.allclose(rh[-1], v0 * T + a * dt * dt * N * (N + 1) / 2, rtol=1e-12, atol=1e-20)


def test_yoshida_position_coefficients_sum():
    """The reference's Yoshida d-coefficients sum to (3 - 2^(1/3))/(2 - 2^(1/3)) ~ 2.3512, not 1:
    with zero fields one step drifts by 2.3512 * v * dt (pusher.py:176-181, 184-209)."""
    v0 = np.array([[1e5, -2e5, 3e4]])
    r0 = np.zeros((1, 3))
    rn, vn = OraclePusher(QM_P, "yoshida").push_batch(r0, v0, const((0, 0, 0)), const((0, 0, 0)), 1e-9)
    cbrt2 = 2.0 ** (1.0 / 3.0)
    assert np.allclose(rn, (3.0 - cbrt2) / (2.0 - cbrt2) * v0 * 1e-9, rtol=1e-14)
    assert np.array_equal(vn, v0)


def _gyro_error(alg, dt, T, qm=QM_P, Bz=1.0, v=1e6):
    N = int(round(T / dt))
    r0 = np.zeros((1, 3))
    v0 = np.array([[v, 0.0, 0.0]])
    rh, _vh, _ = OraclePusher(qm, alg).track_batch(r0, v0, const((0, 0, 0)), const((0, 0, Bz)), N, dt, N)
    w = qm * Bz
    rl = v / w
    exact = np.array([rl * np.sin(w * T), rl * (np.cos(w * T) - 1.0), 0.0])
    return np.linalg.norm(rh[-1, 0] - exact) / rl


def test_rk4_fourth_order_convergence():
    T = 2e-7
    e1 = _gyro_error("rk4", 2e-10, T)
    e2 = _gyro_error("rk4", 1e-10, T)
    assert 14.0 < e1 / e2 < 18.0


def test_rk4_rel_fourth_order_in_1d_rf_field():
    """1-D E = E0 cos(w t + phi), B = 0:  gamma v(t) = gamma v0 + (q E0 / m w)[sin(w t + phi) - sin phi]."""
    qm = OS.q_over_m("electron")
    E0, f, phi = 1e6, 1e8, 0.7
    w = 2 * np.pi * f
    T = 2e-9

    def run(dt):
        N = int(round(T / dt))
        ef = ORF(const((E0, 0.0, 0.0)), w, phi, "cos")
        v0 = np.array([[0.3 * CLIGHT, 0.0, 0.0]])
        _rh, vh, _ = OraclePusher(qm, "rk4_rel").track_batch(np.zeros((1, 3)), v0, ef, const((0, 0, 0)), N, dt, N)
        g0 = 1.0 / np.sqrt(1 - 0.09)
        u_exact = g0 * 0.3 * CLIGHT + qm * E0 / w * (np.sin(w * T + phi) - np.sin(phi))
        v_exact = u_exact / np.sqrt(1 + (u_exact / CLIGHT) ** 2)
        return abs(vh[-1, 0, 0] - v_exact) / CLIGHT

    e1, e2 = run(4e-12), run(2e-12)
    assert e1 < 1e-9
    assert 12.0 < e1 / e2 < 20.0


def test_rk4_rel_reduces_to_relativistic_gyration():
    """Pure B: |v| constant, angular frequency q B / (gamma m)."""
    qm = OS.q_over_m("electron")
    beta = 0.8
    g = 1 / np.sqrt(1 - beta ** 2)
    Bz = 0.01
    w = abs(qm) * Bz / g
    T = 2 * np.pi / w
    N = 4000
    v0 = np.array([[beta * CLIGHT, 0.0, 0.0]])
    _rh, vh, _ = OraclePusher(qm, "rk4_rel").track_batch(np.zeros((1, 3)), v0, const((0, 0, 0)),
                                                         const((0, 0, Bz)), N, T / N, N)
    assert np.allclose(vh[-1, 0], v0[0], rtol=0, atol=1e-8 * CLIGHT)


@pytest.mark.parametrize("alg", ["boris_rel", "vay"])
def test_relativistic_gyro_angle_and_norm(alg):
    """Pure B: |u| conserved to rounding and each step rotates by 2 atan(|q/m| B dt / (2 gamma))."""
    qm = OS.q_over_m("electron")
    beta, Bz, dt, N = 0.9, 0.5, 1e-12, 300
    g = 1 / np.sqrt(1 - beta ** 2)
    v0 = np.array([[beta * CLIGHT, 0.0, 0.0]])
    _rh, vh, _ = OraclePusher(qm, alg).track_batch(np.zeros((1, 3)), v0, const((0, 0, 0)), const((0, 0, Bz)),
                                                   N, dt, 1)
    ang = np.unwrap(np.angle(vh[:-1, 0, 0] + 1j * vh[:-1, 0, 1]))
    dth = np.diff(ang)
    expect = 2 * np.arctan(abs(qm) * Bz * dt / (2 * g))
    assert np.allclose(np.abs(dth), expect, rtol=1e-9)
    assert np.all(np.sign(dth) == np.sign(-qm))          # electrons rotate counter-clockwise about +B
    spd = np.linalg.norm(vh[:, 0], axis=1)
    assert np.allclose(spd, beta * CLIGHT, rtol=1e-13)


def _rotate(u, axis, angle):
    """Rodrigues rotation of the rows of u by `angle` (right-handed) about the unit vectors `axis`."""
    c, s = np.cos(angle)[:, None], np.sin(angle)[:, None]
    return u * c + np.cross(axis, u) * s + axis * np.sum(axis * u, axis=1)[:, None] * (1.0 - c)


def _boris_rel_closed_form(u, E, B, dt, qm):
    """Relativistic Boris as its geometric definition: half E kick, exact rotation of u about -B by
    2 atan(|q/m| |B| dt / (2 gamma_-)), half E kick (Birdsall & Langdon 15-4)."""
    dE = 0.5 * qm * dt * E
    um = u + dE
    gm = np.sqrt(1.0 + np.sum(um * um, axis=1) / CLIGHT ** 2)
    bn = np.linalg.norm(B, axis=1)
    ang = 2.0 * np.arctan(np.abs(qm) * bn * dt / (2.0 * gm))
    axis = -np.sign(qm) * B / np.where(bn > 0, bn, 1.0)[:, None]
    return _rotate(um, axis, ang) + dE


def test_vay_exact_ExB_drift_and_boris_rel_closed_form():
    """Vay (2008) reproduces the E x B drift exactly for any gamma.  Relativistic Boris does not;
    its step equals its geometric definition (half kick, exact rotation, half kick) to rounding."""
    qm = OS.q_over_m("proton")
    B = np.array([0.0, 0.0, 1.0])
    E = np.array([0.0, 0.9 * CLIGHT, 0.0])          # |E| = 0.9 c |B|: E x B / B^2 = 0.9 c along +x
    vd = np.cross(E, B) / np.dot(B, B)
    v0 = vd[None, :]
    outs = {}
    for alg in ("vay", "boris_rel"):
        pu = OraclePusher(qm, alg)
        rn, vn = pu.push_batch(np.zeros((1, 3)), v0, const(E), const(B), 1e-9)
        outs[alg] = vn
    assert np.allclose(outs["vay"], v0, rtol=1e-12, atol=1e-6)
    u1 = _boris_rel_closed_form(K.u_from_v(v0), E[None, :], B[None, :], 1e-9, qm)
    assert np.linalg.norm(outs["boris_rel"] - K.v_from_u(u1)) < 1e-13 * np.linalg.norm(v0)
    assert np.linalg.norm(outs["boris_rel"] - v0) > 1e-6 * np.linalg.norm(v0)   # not the drift


def test_boris_rel_kick_matches_geometric_definition():
    """Random u, E, B over four decades of field strength: the oracle's relativistic Boris step equals
    the half-kick / exact-rotation / half-kick closed form, for both charge signs."""
    rng = np.random.default_rng(12)
    M = 2000
    u = rng.standard_normal((M, 3)) * 0.5 * CLIGHT
    E = rng.standard_normal((M, 3)) * 10.0 ** rng.uniform(4, 8, (M, 1))
    B = rng.standard_normal((M, 3)) * 10.0 ** rng.uniform(-2, 1, (M, 1))
    for qm, dt in ((QM_P, 1e-10), (OS.q_over_m("electron"), 1e-13)):
        got = K.boris_rel_kick_ext(u, E, B, dt, qm)
        want = _boris_rel_closed_form(u, E, B, dt, qm)
        scale = np.linalg.norm(u, axis=1) + np.linalg.norm(0.5 * qm * dt * E, axis=1)
        assert np.max(np.linalg.norm(got - want, axis=1) / scale) < 1e-13


def test_boris_rel_pure_E_and_E_parallel_B_kicks_are_exact():
    """With B = 0, relativistic Boris gives u_N = u_0 + (q/m) E N dt exactly (the two half kicks add to
    one full kick).  With E parallel to B the kick along B is unchanged by the rotation: u_par grows by
    (q/m) E dt per step and |u_perp| is conserved.  A wrong E factor or a dropped half kick fails both."""
    qm, dt, N = OS.q_over_m("electron"), 1e-12, 1000
    rng = np.random.default_rng(13)
    u0 = rng.standard_normal((8, 3)) * 1e8
    E = np.array([[2e5, -1e5, 3e5]])
    u = u0.copy()
    for _ in range(N):
        u = K.boris_rel_kick_ext(u, E, np.zeros((1, 3)), dt, qm)
    want = u0 + qm * E * (N * dt)
    assert np.max(np.linalg.norm(u - want, axis=1) / np.linalg.norm(want, axis=1)) < 1e-12
    # the same through track_batch: v_hist holds v(u) of the staggered scheme at every record
    rh, vh, _ = OraclePusher(qm, "boris_rel").track_batch(np.zeros((8, 3)), K.v_from_u(u0), const(E[0]),
                                                           const((0.0, 0.0, 0.0)), N, dt, N)
    u_back = K.u_from_v(vh[-1])
    assert np.max(np.linalg.norm(u_back - want, axis=1) / np.linalg.norm(want, axis=1)) < 1e-10
    # E parallel to B along z
    Ez, Bz = 4e5, 0.7
    u = u0.copy()
    for _ in range(N):
        u = K.boris_rel_kick_ext(u, np.array([[0.0, 0.0, Ez]]), np.array([[0.0, 0.0, Bz]]), dt, qm)
    scale = np.linalg.norm(u0, axis=1) + abs(qm * Ez * N * dt)
    assert np.max(np.abs(u[:, 2] - (u0[:, 2] + qm * Ez * N * dt)) / scale) < 1e-12
    assert np.max(np.abs(np.hypot(u[:, 0], u[:, 1]) - np.hypot(u0[:, 0], u0[:, 1])) / scale) < 1e-12


def test_vay_rel_speed_growth_quirk():
    """The reference 'vay_rel' uses gamma'^-2 = 1 - u'^2/c^2 and so speeds up in a pure B field:
    per step |v_new|/|v| = gamma_n / gamma_new (> 1); it saturates at the 0.9999 c clamp."""
    ke = 1.0  # MeV
    mp = OS.mass_mev("proton")
    g = 1 + ke / mp
    beta = np.sqrt(1 - 1 / g ** 2)
    v0 = np.array([[beta * CLIGHT, 0.0, 0.0]])
    _r, v1 = OraclePusher(QM_P, "vay_rel").push_batch(np.zeros((1, 3)), v0, const((0, 0, 0)), const((0, 0, 1.0)),
                                                       1e-11)
    ratio = np.linalg.norm(v1) / np.linalg.norm(v0)
    assert 1.0020 < ratio < 1.0023
    rh, vh, _ = OraclePusher(QM_P, "vay_rel").track_batch(np.zeros((1, 3)), v0, const((0, 0, 0)),
                                                           const((0, 0, 1.0)), 800, 1e-11, 800)
    assert np.linalg.norm(vh[-1, 0]) == pytest.approx(0.9999 * CLIGHT, rel=1e-12)


# --------------------------------------------------------------------------- gather


def _grid(rng, n, nonuniform=False):
    g = np.sort(rng.uniform(-1, 1, n)) if nonuniform else np.linspace(-1, 1, n)
    return g


@pytest.mark.parametrize("nonuniform", [False, True])
def test_trilinear_exact_on_multilinear_fields(nonuniform):
    rng = np.random.default_rng(3)
    gx, gy, gz = _grid(rng, 7, nonuniform), _grid(rng, 5, nonuniform), _grid(rng, 9, nonuniform)
    co = rng.standard_normal(8)
    f = lambda x, y, z: (co[0] + co[1] * x + co[2] * y + co[3] * z + co[4] * x * y + co[5] * x * z
                         + co[6] * y * z + co[7] * x * y * z)
    X, Y, Z = np.meshgrid(gx, gy, gz, indexing="ij")
    vals = f(X, Y, Z)
    pts = np.column_stack([rng.uniform(gx[0], gx[-1], 5000), rng.uniform(gy[0], gy[-1], 5000),
                           rng.uniform(gz[0], gz[-1], 5000)])
    got = interp_nd((gx, gy, gz), vals, pts)
    assert np.max(np.abs(got - f(*pts.T))) < 1e-13


def test_trilinear_matches_scipy_routines_and_fill():
    rng = np.random.default_rng(4)
    gx, gy, gz = np.linspace(0, 1, 11), np.linspace(-2, 2, 8), np.linspace(3, 4, 13)
    vals = rng.standard_normal((11, 8, 13))
    pts = np.column_stack([rng.uniform(-0.1, 1.1, 20000), rng.uniform(-2.2, 2.2, 20000), rng.uniform(2.9, 4.1, 20000)])
    ours = interp_nd((gx, gy, gz), vals, pts, 0.0)
    rgi = RegularGridInterpolator((gx, gy, gz), vals, bounds_error=False, fill_value=0.0)(pts)
    cm = interp_coordinate_mapper((gx, gy, gz), vals, pts, 0.0)
    assert np.max(np.abs(ours - rgi)) < 1e-14
    assert np.max(np.abs(ours - cm)) < 1e-14
    outside = (pts[:, 0] < 0) | (pts[:, 0] > 1) | (pts[:, 1] < -2) | (pts[:, 1] > 2) | (pts[:, 2] < 3) | (pts[:, 2] > 4)
    assert np.all(ours[outside] == 0.0) and outside.sum() > 1000
    # upper faces are inside and give the face value
    face = np.array([[1.0, 2.0, 4.0], [0.0, -2.0, 3.0], [1.0, 0.5, 3.5]])
    got = interp_nd((gx, gy, gz), vals, face)
    assert got[0] == vals[-1, -1, -1] and got[1] == vals[0, 0, 0]
    assert np.isclose(got[2], RegularGridInterpolator((gx, gy, gz), vals)(face[2:])[0], rtol=1e-14)


def test_periodic_z_wrap_matches_shifted_evaluation():
    rng = np.random.default_rng(5)
    nz, dz = 16, 1e-3
    grid = {"x": np.linspace(-3e-3, 3e-3, 7), "y": np.linspace(-3e-3, 3e-3, 7), "z": np.arange(nz) * dz}
    X, Y, Z = np.meshgrid(grid["x"], grid["y"], grid["z"], indexing="ij")
    L = nz * dz
    vals = {"x": X * np.cos(2 * np.pi * Z / L), "y": Y * np.sin(2 * np.pi * Z / L), "z": 1 + 0 * Z}
    f = OPeriodicZ(grid, vals)
    pts = np.column_stack([rng.uniform(-3e-3, 3e-3, 2000), rng.uniform(-3e-3, 3e-3, 2000), rng.uniform(0, L, 2000)])
    base = f(pts)
    for k in (-3, 1, 7, 250):
        sh = pts.copy()
        sh[:, 2] += k * L
        assert np.max(np.abs(f(sh) - base)) < 1e-9 * np.max(np.abs(base))
    # inside [z0, z_last] the periodic map equals the plain map
    inner = pts[pts[:, 2] <= grid["z"][-1]]
    plain = OField.from_arrays(grid, vals)
    assert np.array_equal(f(inner), plain(inner))


@pytest.mark.parametrize("z0", [0.0, 0.0137, -0.25])
def test_periodic_wrap_cell_interpolates_last_and_first_slices(z0):
    """The wrap cell (z_last, z0 + L) joins slice nz-1 to slice 0: at z = z_last + f dz the periodic map
    is (1 - f) S_{nz-1}(x, y) + f S_0(x, y), with S_k the bilinear interpolation of slice k (scipy's
    RegularGridInterpolator on the 2-D slice).  The same holds below z0 (z0 - f dz wraps to weight
    1 - f) and k periods away, and the wrap is about z0, not about 0.  Random node values, so no
    smooth structure hides a wrong slice."""
    rng = np.random.default_rng(17)
    nz, dz = 12, 1e-3
    grid = {"x": np.linspace(-3e-3, 3e-3, 6), "y": np.linspace(-2e-3, 4e-3, 5), "z": z0 + np.arange(nz) * dz}
    vals = {c: rng.standard_normal((6, 5, nz)) for c in "xyz"}
    f = OPeriodicZ(grid, vals)
    L = nz * dz
    M = 3000
    x, y = rng.uniform(-3e-3, 3e-3, M), rng.uniform(-2e-3, 4e-3, M)
    fr = rng.uniform(0.0, 1.0, M)
    xy = np.column_stack([x, y])
    want = np.column_stack([
        (1.0 - fr) * RegularGridInterpolator((grid["x"], grid["y"]), vals[c][:, :, -1])(xy)
        + fr * RegularGridInterpolator((grid["x"], grid["y"]), vals[c][:, :, 0])(xy) for c in "xyz"])
    z_last = grid["z"][-1]
    for shift in (0, 1, -2, 9):
        got = f(np.column_stack([x, y, z_last + fr * dz + shift * L]))
        assert np.max(np.abs(got - want)) < 1e-9 * np.max(np.abs(want))
    below = f(np.column_stack([x, y, z0 - (1.0 - fr) * dz]))
    assert np.max(np.abs(below - want)) < 1e-9 * np.max(np.abs(want))
    # node planes: z = z0 + k L is slice 0 exactly, z = z_last + k L is slice nz-1
    for k in (0, 3, -1):
        at0 = f(np.column_stack([x, y, np.full(M, z0 + k * L)]))
        w0 = np.column_stack([RegularGridInterpolator((grid["x"], grid["y"]), vals[c][:, :, 0])(xy) for c in "xyz"])
        assert np.max(np.abs(at0 - w0)) < 1e-9 * np.max(np.abs(w0))
