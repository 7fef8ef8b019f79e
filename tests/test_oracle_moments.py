"""The oracle's restatement of ParticleDistribution's moments (particles.py:685-843),
pinned by identities and closed forms that a dropped term or a wrong sign breaks."""
import numpy as np

from oracle import moments as OM
from oracle.constants import CLIGHT


def _beam(n=20000, seed=3, alpha=-1.3, beta=2.0, eps=3e-6):
    """x, x' drawn on a Twiss ellipse (alpha, beta, eps); y uncorrelated; z along the motion."""
    rng = np.random.default_rng(seed)
    u, w = rng.standard_normal(n), rng.standard_normal(n)
    x = np.sqrt(eps * beta) * u
    xp = np.sqrt(eps / beta) * (w - alpha * u)
    y = 1e-3 * rng.standard_normal(n)
    yp = 2e-3 * rng.standard_normal(n)
    pz = 0.05 * (1 + 1e-3 * rng.standard_normal(n))
    p = np.column_stack([xp * pz, yp * pz, pz])
    r = np.column_stack([x, y, 0.01 * rng.standard_normal(n)])
    return r, p


def test_twiss_closed_form_and_identity():
    r, p = _beam()
    v = OM.physical_v_from_p(p)
    m = OM.moments(r, p, v, 938.272)
    ax, bx, gx, ay, by, gy = OM.twiss(m)
    ex = OM.emittances(m, normalized=False)[0]
    # alpha = -<x x'>/eps and beta gamma - alpha^2 = 1 (exact for rms Twiss parameters)
    np.testing.assert_allclose(ax, -m["xxp"] / ex, rtol=1e-9)
    np.testing.assert_allclose(bx * gx - ax * ax, 1.0, rtol=1e-12)
    np.testing.assert_allclose(by * gy - ay * ay, 1.0, rtol=1e-12)
    assert ax < -1.0            # drawn with alpha = -1.3: a diverging beam has <x x'> > 0
    np.testing.assert_allclose([ax, bx, ex], [-1.3, 2.0, 3e-6], rtol=0.05)


def test_emittance_invariant_under_thin_lens_and_drift():
    """x' -> x' + k x and x -> x + L x' are symplectic: the rms emittance does not change."""
    r, p = _beam(seed=5)
    v = OM.physical_v_from_p(p)
    e0 = OM.emittances(OM.moments(r, p, v, 938.272), normalized=False)[:2]
    xp = v[:, 0] / v[:, 2]
    for k, L in ((-0.7, 0.0), (0.0, 1.9), (2.5, -0.4)):
        xp2 = xp + k * r[:, 0]
        r2 = r.copy()
        r2[:, 0] = r[:, 0] + L * xp2
        v2 = v.copy()
        v2[:, 0] = xp2 * v[:, 2]
        e = OM.emittances(OM.moments(r2, p, v2, 938.272), normalized=False)[:2]
        np.testing.assert_allclose(e, e0, rtol=1e-8)


def test_energy_and_conversions():
    rng = np.random.default_rng(7)
    p = rng.standard_normal((1000, 3)) * 0.3
    m = 938.272
    g = np.sqrt(1.0 + np.sum(p * p, axis=1))
    # m (gamma - 1) in the cancellation-free form m p^2 / (gamma + 1); the reference's form
    # sqrt((p m)^2 + m^2) - m loses digits when p is small, hence the absolute tolerance
    np.testing.assert_allclose(OM.ekin_mev(p, m), m * (g * g - 1.0) / (g + 1.0), rtol=1e-10, atol=1e-11)
    pz2 = p[:, 2] ** 2
    np.testing.assert_allclose(OM.ekin_mev(p, m, z_energy=True), m * pz2 / (np.sqrt(1 + pz2) + 1.0), rtol=1e-10,
                               atol=1e-11)
    v = OM.physical_v_from_p(p)
    np.testing.assert_allclose(OM.physical_p_from_v(v), p, rtol=1e-12, atol=1e-15)
    assert np.all(np.linalg.norm(v, axis=1) < CLIGHT)
    # the reference's per-component conversion agrees with the vector form along one axis only
    pa = np.column_stack([np.zeros(10), np.zeros(10), np.linspace(0.01, 3, 10)])
    np.testing.assert_allclose(OM.reference_v_from_p(pa), OM.physical_v_from_p(pa), rtol=1e-15)
    vr = OM.reference_v_from_p(p)
    assert np.max(np.abs(vr - v) / np.linalg.norm(v, axis=1)[:, None]) > 1e-3
    np.testing.assert_allclose(OM.reference_p_from_v(OM.reference_v_from_p(p)), p, rtol=1e-9, atol=1e-14)


def test_edges_and_angle_guard():
    r = np.array([[0.0, 1.0, 2.0], [2.0, -1.0, 4.0], [4.0, 0.0, 0.0]])
    p = np.array([[0.1, 0.0, 0.0], [0.0, 0.1, 1e-20], [0.0, 0.0, 0.2]])
    v = OM.physical_v_from_p(p)
    m = OM.moments(r, p, v, 1.0)
    np.testing.assert_array_equal(OM.beam_edges(m, "full"), [0.0, 4.0, -1.0, 1.0, 0.0, 4.0])
    np.testing.assert_allclose(OM.beam_edges(m, "1rms")[:2], [2.0 - np.sqrt(8 / 3), 2.0 + np.sqrt(8 / 3)])
    a = OM.angles(v)
    assert a[0, 0] == v[0, 0] / 1e-10 and a[2, 2] == 1.0      # |v_z| < 1e-10 -> 1e-10
