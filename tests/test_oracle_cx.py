"""Pins the oracle's charge-exchange integrals.

* or_cx_rate_ref restates MK_source.cpp:38-103; it is compared with the reference program's own
  output (tests/golden/mk_source_int1.json, made by tests/golden/make_mk_source_golden.py from the
  reference compiled under oracle/_ref), and for sigma = const with the closed form of the mean
  relative speed of a shifted Maxwellian, I/sigma = c [exp(-w^2)/sqrt(pi) + (w + 1/(2w)) erf w],
  w = U/c, converging at second order.
* or_cx_source (the exact source of the surrogate) is checked against an independent 3-D
  Cartesian quadrature of the relative-velocity integrals, a 6-D Monte Carlo of the energy
  exchange, its bilinearity in the densities and Galilean invariance.  The Lindsay-Stebbings
  sigma is the published fit (J. Geophys. Res. 110, A12213, 2005); no reference file pins it, so
  its absolute value is "parity unpinned" with respect to TITAN, which uses (1 - a2 ln u)^2.
"""
import json
import math
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def test_mk_source_golden(O):
    with open(os.path.join(HERE, "golden", "mk_source_int1.json")) as f:
        g = json.load(f)
    for p in g["points"]:
        got = O.cx_rate_ref(p["U"], p["cp"], g["a_2"], g["N1"], g["N2"])
        assert abs(got - p["I"]) <= 1e-12 * abs(p["I"]), p


@pytest.mark.parametrize("U,c", [(0.5, 1.0), (3.0, 1.0), (10.0, 2.0), (0.01, 5.0), (100.0, 3.0)])
def test_constant_sigma_closed_form(O, U, c):
    w = U / c
    exact = c * (math.exp(-w * w) / math.sqrt(math.pi) + (w + 1 / (2 * w)) * math.erf(w))
    e100 = O.cx_rate_ref(U, c, 0.0, 100, 100) - exact
    e400 = O.cx_rate_ref(U, c, 0.0, 400, 400) - exact
    assert abs(e100) < 5e-5 * exact
    assert 14.0 < e100 / e400 < 18.0   # midpoint rule: second order


def test_sigma_ls_published_fit(O):
    # E = 1 keV: (4.15 - 0)^2 (1 - e^-67.3)^4.5 1e-16 cm^2 = 1.72225e-15 cm^2
    e_per_g2 = 0.5 * 1.67262192369e-24 * (1.03804e6 ** 2) / 1.602176634e-9
    g = math.sqrt(1.0 / e_per_g2)
    nl = 0.06 * 4.2113 * 1.495978707e13
    assert abs(O.sigma_ls(g) / nl - 1.72225e-15) < 1e-19
    # 100 keV: steep fall-off factor (1 - exp(-0.673))^4.5
    g100 = math.sqrt(100.0 / e_per_g2)
    exp100 = (4.15 - 0.531 * math.log(100.0)) ** 2 * (1 - math.exp(-0.673)) ** 4.5 * 1e-16
    assert abs(O.sigma_ls(g100) / nl / exp100 - 1) < 1e-13


def _sigma_np(O, g):
    e_per_g2 = 0.5 * 1.67262192369e-24 * (1.03804e6 ** 2) / 1.602176634e-9
    nl = 0.06 * 4.2113 * 1.495978707e13
    E = np.maximum(e_per_g2 * g * g, 1e-12)
    return (4.15 - 0.531 * np.log(E)) ** 2 * (1 - np.exp(-67.3 / E)) ** 4.5 * 1e-16 * nl


CASES = [
    (1.0, 1.2, [0.5, -2.0, 0.3], 0.5, 0.9, [-2.5, 0.2, 0.1]),
    (0.7, 4.0, [3.0, 0.0, 0.0], 2.0, 1.5, [-2.0, 0.5, -0.4]),
    (3.0, 0.3, [-1.0, 1.0, 1.0], 1.0, 2.5, [-1.0, 1.0, 1.2]),
]


@pytest.mark.parametrize("case", CASES)
def test_cx_source_vs_cartesian_quadrature(O, case):
    n_p, cp2, vp, n_h, ch2, vh = case
    vp, vh = np.array(vp), np.array(vh)
    out = O.cx_source(n_p, cp2, vp, n_h, ch2, vh, nodes=200)
    dv = vp - vh
    c2 = cp2 + ch2
    c = math.sqrt(c2)
    L, m = 6.0 * c, 161
    ax = np.linspace(-L, L, m)
    h = ax[1] - ax[0]
    X, Y, Z = np.meshgrid(ax, ax, ax, indexing="ij")
    gx, gy, gz = X + dv[0], Y + dv[1], Z + dv[2]       # g = dV + deviation
    gm = np.sqrt(gx * gx + gy * gy + gz * gz)
    f = np.exp(-(X * X + Y * Y + Z * Z) / c2) / (math.pi ** 1.5 * c2 * c)
    ws = f * gm * _sigma_np(O, gm) * h ** 3
    J0 = ws.sum()
    Jv = np.array([(ws * gx).sum(), (ws * gy).sum(), (ws * gz).sum()])
    J2 = (ws * gm * gm).sum()
    nn = n_p * n_h
    assert abs(out[0] - nn * J0) < 2e-4 * nn * J0
    np.testing.assert_allclose(out[1:4], -nn * Jv, rtol=0, atol=3e-4 * nn * np.abs(Jv).max())
    kappa = (cp2 - ch2) / c2
    U = np.linalg.norm(dv)
    e = dv / U
    J1 = Jv @ e
    SE = -0.5 * nn * ((vp + vh) @ e * J1 + kappa * (J2 - U * J1))
    assert abs(out[4] - SE) < 3e-4 * abs(SE) + 1e-12


def test_cx_energy_monte_carlo(O):
    """6-D Monte Carlo of S_E = n_p n_H <g sigma(g) (|v_H|^2 - |v_p|^2)/2> over independent
    Maxwellians: pins the conditional-mean (kappa) term of the closed reduction."""
    n_p, cp2, vp, n_h, ch2, vh = 1.0, 3.0, np.array([1.0, -0.5, 0.2]), 1.0, 0.5, np.array([-1.5, 0.3, 0.0])
    rng = np.random.default_rng(5)
    N = 4_000_000
    wp = rng.normal(0, math.sqrt(cp2 / 2), (N, 3)) + vp
    wh = rng.normal(0, math.sqrt(ch2 / 2), (N, 3)) + vh
    g = np.linalg.norm(wp - wh, axis=1)
    r = g * _sigma_np(O, g)
    SE_mc = (r * 0.5 * ((wh * wh).sum(1) - (wp * wp).sum(1))).mean()
    SM_mc = (r[:, None] * (wh - wp)).mean(axis=0)
    err = (r * 0.5 * ((wh * wh).sum(1) - (wp * wp).sum(1))).std() / math.sqrt(N)
    out = O.cx_source(n_p, cp2, vp, n_h, ch2, vh, nodes=200)
    assert abs(out[4] - SE_mc) < 5 * err
    assert np.abs(out[1:4] - SM_mc).max() < 5 * (r * np.abs(wh - wp).max(1)).std() / math.sqrt(N)


def test_cx_bilinear_and_galilean(O):
    vp, vh = np.array([0.3, 1.0, -2.0]), np.array([-2.5, 0.4, 0.0])
    a = O.cx_source(1.0, 2.0, vp, 1.0, 1.0, vh)
    b = O.cx_source(3.0, 2.0, vp, 0.5, 1.0, vh)
    np.testing.assert_allclose(b, 1.5 * a, rtol=1e-14)
    shift = np.array([5.0, -1.0, 2.0])
    c = O.cx_source(1.0, 2.0, vp + shift, 1.0, 1.0, vh + shift)
    np.testing.assert_allclose(c[:4], a[:4], rtol=1e-13)
    # plasma-frame energy exchange is frame independent: S_E - V_p . S_m
    assert abs((c[4] - (vp + shift) @ c[1:4]) - (a[4] - vp @ a[1:4])) < 1e-12 * abs(a[4])
    # momentum is anti-parallel to V_p - V_H
    dv = vp - vh
    cross = np.cross(a[1:4], dv)
    assert np.abs(cross).max() < 1e-13 * np.abs(a[1:4]).max() * np.linalg.norm(dv)
    assert a[1:4] @ dv < 0
