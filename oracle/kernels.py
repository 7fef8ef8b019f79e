"""Integrator kernels of the reference, restated in NumPy FP64.  Test infrastructure only.

Every kernel works on (M, 3) float64 arrays (one row per particle) and follows
the reference's operation order operation by operation, so that NumPy's
per-element IEEE arithmetic gives a bitwise-defined result:

* ``a * b * c`` in Python is ``(a * b) * c``; ``x + y + z`` is ``(x + y) + z``.
* ``np.cross`` and the reference's ``cross_product`` both compute
  ``(a1*b2 - a2*b1, a2*b0 - a0*b2, a0*b1 - a1*b0)`` (pusher.py:99-106).
* ``np.sum(a, axis=1)`` over 3 columns is ``(a0 + a1) + a2`` (checked in
  tests/test_oracle_semantics.py).

The ``sq`` argument selects how the reference's ``x ** 2`` is evaluated:

* ``sq_mul`` (default, *canonical form*): ``x * x``.  This is what the Numba
  build of the reference compiles ``x**2`` to and what ``np.sum(v**2)`` does.
* ``sq_pow`` (*as-shipped form*): libm ``pow(x, 2.0)``, which is what a NumPy
  float64 scalar's ``**`` calls in the reference's no-numba per-particle path
  (pusher.py:755-775).  It differs from ``x*x`` by one ulp in ~0.1 % of cases.

Kernels named ``*_ext`` are engine extensions with no reference symbol
('boris_rel' relativistic Boris, textbook 'vay' of Vay 2008); they are pinned by
closed forms only.
"""
import math

import numpy as np

from .constants import CLIGHT

C2 = CLIGHT ** 2          # pusher.py:90 ``CLIGHT ** 2``
VCLAMP = 0.9999 * CLIGHT  # pusher.py:276

# Yoshida coefficients, pusher.py:176-181 (same Python expressions).
YOSHIDA_C1 = 1.0 / (2.0 - 2.0 ** (1.0 / 3.0))
YOSHIDA_C2 = -2.0 ** (1.0 / 3.0) / (2.0 - 2.0 ** (1.0 / 3.0))
YOSHIDA_C3 = YOSHIDA_C1
YOSHIDA_D1 = 2.0 * YOSHIDA_C1
YOSHIDA_D2 = -(2.0 ** (1.0 / 3.0) + 1.0) / (2.0 - 2.0 ** (1.0 / 3.0))
YOSHIDA_D3 = YOSHIDA_D1


def sq_mul(x):
    return x * x


_pow2 = np.frompyfunc(lambda a: math.pow(a, 2.0), 1, 1)


def sq_pow(x):
    return np.asarray(_pow2(x), dtype=np.float64)


def cross(a, b):
    """pusher.py:99-106 (component order of cross_product)."""
    out = np.empty(np.broadcast_shapes(a.shape, b.shape))
    out[:, 0] = a[:, 1] * b[:, 2] - a[:, 2] * b[:, 1]
    out[:, 1] = a[:, 2] * b[:, 0] - a[:, 0] * b[:, 2]
    out[:, 2] = a[:, 0] * b[:, 1] - a[:, 1] * b[:, 0]
    return out


def norm2(a, sq=sq_mul):
    return (sq(a[:, 0]) + sq(a[:, 1])) + sq(a[:, 2])


def dot(a, b):
    return (a[:, 0] * b[:, 0] + a[:, 1] * b[:, 1]) + a[:, 2] * b[:, 2]


def _col(x):
    """Broadcast a scalar or per-particle (M,) array against (M, 3)."""
    x = np.asarray(x, dtype=np.float64)
    return x[:, None] if x.ndim == 1 else x


# ---------------------------------------------------------------------------
# Reference kernels
# ---------------------------------------------------------------------------

def relativistic_gamma(v, sq=sq_mul):
    """pusher.py:74-96: gamma from |v| with beta^2 clamped to 0.9999."""
    beta_sq = norm2(v, sq) / C2
    beta_sq = np.where(beta_sq >= 0.9999, 0.9999, beta_sq)
    return 1.0 / np.sqrt(1.0 - beta_sq)


def lorentz_accel(v, E, B, qm):
    """leapfrog_dbetagamma_dt / rk4_dbetagamma_dt, pusher.py:113-121, 168-172."""
    return qm * (E + cross(v, B))


def leapfrog_step(r, v, E, B, dt, qm):
    """push_batch 'leapfrog' NumPy path, pusher.py:742-745 (push(): :539-542)."""
    dv = lorentz_accel(v, E, B, qm)
    v_new = v + dv * dt
    return r + v_new * dt, v_new


def boris_kick(v, E, B, dt, qm, sq=sq_mul):
    """boris_push_single, pusher.py:124-165 (non-relativistic Boris)."""
    half = 0.5 * qm
    v_minus = v + half * E * dt
    t = half * B * dt
    t_mag_sq = norm2(t, sq)
    s = 2.0 * t / (1.0 + t_mag_sq)[:, None]
    v_prime = v_minus + cross(v_minus, t)
    v_plus = v_minus + cross(v_prime, s)
    return v_plus + half * E * dt


def boris_step(r, v, E, B, dt, qm, sq=sq_mul):
    """push_batch 'boris', pusher.py:747-760: kick then r + v_new*dt."""
    v_new = boris_kick(v, E, B, dt, qm, sq)
    return r + v_new * dt, v_new


def vay_rel_kick(v, E, B, dt, qm, sq=sq_mul):
    """vay_push_single, pusher.py:216-279, including its gamma formula and clamps."""
    gamma_n = relativistic_gamma(v, sq)
    u_n = gamma_n[:, None] * v
    half = 0.5 * qm
    tau = half * B * dt
    u_prime = u_n + qm * E * dt
    u_prime_mag = np.sqrt(norm2(u_prime, sq))
    tau_mag = np.sqrt(norm2(tau, sq))
    u_prime_dot_tau = dot(u_prime, tau)
    g_inv_sq = 1.0 - sq(u_prime_mag / CLIGHT)
    g_inv_sq = np.where(g_inv_sq < 1e-10, 1e-10, g_inv_sq)
    tau_mag_sq = sq(tau_mag)
    sigma = g_inv_sq - tau_mag_sq
    gamma_new = np.sqrt(0.5 * (sigma + np.sqrt(sq(sigma) + 4.0 * (tau_mag_sq + sq(u_prime_dot_tau / CLIGHT)))))
    gcol = gamma_new[:, None]
    t = tau / gcol
    s = 2.0 * t / (1.0 + tau_mag_sq / sq(gamma_new))[:, None]
    u_star = u_prime + cross(u_prime, t)
    u_new = u_prime + cross(u_star, s)
    v_new = u_new / gcol
    return clamp_speed(v_new, sq)


def clamp_speed(v, sq=sq_mul):
    """|v| >= 0.9999 c  ->  v * (0.9999 c / |v|)   (pusher.py:275-277, 886-890)."""
    v_mag = np.sqrt(norm2(v, sq))
    over = v_mag >= VCLAMP
    if np.any(over):
        v = v.copy()
        v[over] = v[over] * (VCLAMP / v_mag[over])[:, None]
    return v


def vay_rel_step(r, v, E, B, dt, qm, sq=sq_mul):
    """push_batch 'vay_rel', pusher.py:762-775."""
    v_new = vay_rel_kick(v, E, B, dt, qm, sq)
    return r + v_new * dt, v_new


def rk4_rel_deriv(v, E, B, qm, sq=sq_mul):
    """rk4_rel_dbetagamma_dt(_batch_numpy), pusher.py:282-298, 408-423.

    The batch NumPy form clips beta^2 to [0, 0.9999] and squares with array
    ``**2``; the single-particle form uses relativistic_gamma.  Both reduce to
    the same operations in the canonical form.
    """
    gamma = relativistic_gamma(v, sq)
    force = qm * (E + cross(v, B))
    v_dot_f = dot(v, force)
    correction = (v_dot_f / (gamma * C2))[:, None] * v
    return force / gamma[:, None] - correction


def rk4_stage_times():
    """Stage time offsets (in units of dt) of _rk4_step_batch, pusher.py:818-868."""
    return (0.0, 0.5, 0.5, 1.0)


def rk4_step(r, v, field_eval, dt, qm, relativistic, sq=sq_mul):
    """_rk4_step_batch, pusher.py:803-892.

    ``field_eval(points, stage)`` returns (E, B) at the stage positions; stage is
    0..3 so a time-aware caller can place the stage in time.
    """
    dt = _col(dt)
    if relativistic:
        deriv = lambda vv, ee, bb: rk4_rel_deriv(vv, ee, bb, qm, sq)
    else:
        deriv = lambda vv, ee, bb: lorentz_accel(vv, ee, bb, qm)
    E1, B1 = field_eval(r, 0)
    k1_v = deriv(v, E1, B1)
    k1_r = v.copy()
    r2 = r + 0.5 * dt * k1_r
    v2 = v + 0.5 * dt * k1_v
    E2, B2 = field_eval(r2, 1)
    k2_v = deriv(v2, E2, B2)
    k2_r = v2.copy()
    r3 = r + 0.5 * dt * k2_r
    v3 = v + 0.5 * dt * k2_v
    E3, B3 = field_eval(r3, 2)
    k3_v = deriv(v3, E3, B3)
    k3_r = v3.copy()
    r4 = r + dt * k3_r
    v4 = v + dt * k3_v
    E4, B4 = field_eval(r4, 3)
    k4_v = deriv(v4, E4, B4)
    k4_r = v4.copy()
    r_new = r + (dt / 6.0) * (k1_r + 2.0 * k2_r + 2.0 * k3_r + k4_r)
    v_new = v + (dt / 6.0) * (k1_v + 2.0 * k2_v + 2.0 * k3_v + k4_v)
    if relativistic:
        v_new = clamp_speed(v_new, sq_mul)  # array ``**2`` is an exact square
    return r_new, v_new


def yoshida_step(r, v, E, B, dt, qm, relativistic, sq=sq_mul):
    """yoshida(_rel)_push_single, pusher.py:184-209, 332-362 (fields frozen at step start)."""
    if relativistic:
        acc = lambda vv: rk4_rel_deriv(vv, E, B, qm, sq)
    else:
        acc = lambda vv: lorentz_accel(vv, E, B, qm)
    r_t = r.copy()
    v_t = v.copy()
    for c, d in ((YOSHIDA_C1, YOSHIDA_D1), (YOSHIDA_C2, YOSHIDA_D2), (YOSHIDA_C3, YOSHIDA_D3)):
        v_t = v_t + c * dt * acc(v_t)
        r_t = r_t + d * dt * v_t
    if relativistic:
        v_t = clamp_speed(v_t, sq)
    return r_t, v_t


# ---------------------------------------------------------------------------
# Engine extensions (no reference symbol): relativistic Boris and textbook Vay.
# State is u = gamma * v.
# ---------------------------------------------------------------------------

def gamma_from_u(u):
    return np.sqrt(1.0 + norm2(u) / C2)


def u_from_v(v):
    """u = gamma(v) v with the reference's clamped gamma (pusher.py:74-96)."""
    return relativistic_gamma(v)[:, None] * v


def v_from_u(u):
    return u / gamma_from_u(u)[:, None]


def boris_rel_kick_ext(u, E, B, dt, qm):
    """Relativistic Boris (Boris 1970; Birdsall & Langdon 15-4) on u = gamma v."""
    half = 0.5 * qm
    dE = half * E * dt
    u_minus = u + dE
    g_minus = np.sqrt(1.0 + norm2(u_minus) / C2)
    t = half * B * dt / g_minus[:, None]
    s = 2.0 * t / (1.0 + norm2(t))[:, None]
    u_prime = u_minus + cross(u_minus, t)
    u_plus = u_minus + cross(u_prime, s)
    return u_plus + dE


def vay_kick_ext(u, E, B, dt, qm):
    """Textbook Vay pusher (J.-L. Vay, Phys. Plasmas 15, 056701 (2008), eqs. 11-13)."""
    half = 0.5 * qm
    g = gamma_from_u(u)
    v = u / g[:, None]
    tau = half * B * dt
    u_prime = u + qm * E * dt + cross(v, tau)
    tau2 = norm2(tau)
    u_star = dot(u_prime, tau) / CLIGHT
    sigma = (1.0 + norm2(u_prime) / C2) - tau2
    g_new = np.sqrt(0.5 * (sigma + np.sqrt(sigma * sigma + 4.0 * (tau2 + u_star * u_star))))
    t = tau / g_new[:, None]
    s = 1.0 / (1.0 + norm2(t))
    u_new = s[:, None] * (u_prime + dot(u_prime, t)[:, None] * t + cross(u_prime, t))
    return u_new
