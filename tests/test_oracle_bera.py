"""Pins of the oracle's Calc_sourse_MF_Bera restatement (Setka_physics.cpp:746-1258, plasma part):
the momentum and energy the protons exchange with hydrogen fluids by charge exchange.

Checks that a dropped term, a wrong sign or a wrong operand breaks:
- no atoms (rho_H = 0) -> exactly no source;
- thermal and mechanical equilibrium (same V, same T) -> exactly no source; the hydrogen
  temperature is set to the proton temperature of Plasma_components_1 with rho_He = 0,
  T_p = (1 + m_e/m_p) p / rho, so that split is checked too;
- two identical fluids give exactly the source of one fluid of twice the density;
- cold beams (T -> 0): the textbook transfer rates sigma(|dV|) n_p n_H |dV| (V_H - V_p) and
  sigma n_p n_H |dV| (V_H^2 - V_p^2) / 2, with n_p = rho / (1 + m_e/m_p) and
  sigma = (1 - a_2 ln |dV|)^2;
- rotations: the momentum source rotates with the velocities, the energy source does not change;
- a Galilean boost by W: the momentum source is unchanged and E -> E + W . m.
"""
import numpy as np
import pytest

ME_MP = 1.0 / 1836.15
A2, NH, KN = 0.102046, 3.0, 40.0906
DDP = NH / KN


def _state(rho=0.8, v=(0.3, -0.2, 0.1), p=0.5, b=(0.1, 0.2, -0.3)):
    return np.array([rho, *v, p, *b], np.float64)


def _src(O, s, h):
    return O.bera_source(s[None, :], np.asarray(h, np.float64)[None, :, :])[0]


def test_no_atoms_no_source(O):
    s = _state()
    h = [[0.0, 0.2, 1.0, 0.0, 0.0], [0.0, 0.0, -2.0, 0.5, 0.0]]
    np.testing.assert_array_equal(_src(O, s, h), np.zeros(5))


def test_equilibrium_no_source(O):
    s = _state(rho=1.3, v=(0.4, 0.2, -0.7), p=0.9)
    Tp = 8.0 * (1.0 + ME_MP) * s[4] / (8.0 * s[0])     # Plasma_components_1, rho_He = 0
    h = [[1.0, Tp / 2.0, *s[1:4]], [0.25, 0.25 * Tp / 2.0, *s[1:4]]]
    out = _src(O, s, h)
    np.testing.assert_array_equal(out, np.zeros(5))
    # a proton temperature off by the electron share would not balance
    h_bad = [[1.0, s[4] / s[0] / 2.0, *s[1:4]]]
    assert np.abs(_src(O, s, h_bad)[4]) > 0.0


def test_additive_in_fluids(O):
    s = _state()
    one = [0.7, 0.3, -1.5, 0.4, 0.2]
    two = _src(O, s, [one, one])
    dbl = _src(O, s, [[2 * one[0], 2 * one[1], *one[2:]]])
    np.testing.assert_array_equal(two, dbl)
    assert np.abs(two[1:]).min() > 0.0


@pytest.mark.parametrize("vh", [(-2.0, 0.5, 0.3), (1.5, -0.5, 2.0)])
def test_cold_beams(O, vh):
    rho, vp = 1.1, np.array([0.4, -0.3, 0.2])
    s = _state(rho=rho, v=vp, p=0.0)            # T_p = 0 -> clipped to 1e-8
    rho_h = 0.6
    h = [[rho_h, 0.0, *vh]]                      # T_H = 0
    out = _src(O, s, h)
    vh = np.array(vh)
    du = np.linalg.norm(vh - vp)
    n_p = rho / (1.0 + ME_MP)
    rate = (1.0 - A2 * np.log(du)) ** 2 * n_p * rho_h * du
    assert out[0] == 0.0
    np.testing.assert_allclose(out[1:4], DDP * rate * (vh - vp), rtol=1e-6)
    np.testing.assert_allclose(out[4], DDP * rate * 0.5 * (vh @ vh - vp @ vp), rtol=1e-6)


def _rot(a, b, c):
    ca, sa, cb, sb, cc, sc = np.cos(a), np.sin(a), np.cos(b), np.sin(b), np.cos(c), np.sin(c)
    Rz = np.array([[ca, -sa, 0], [sa, ca, 0], [0, 0, 1]])
    Ry = np.array([[cb, 0, sb], [0, 1, 0], [-sb, 0, cb]])
    Rx = np.array([[1, 0, 0], [0, cc, -sc], [0, sc, cc]])
    return Rz @ Ry @ Rx


def test_rotation_covariance(O):
    s = _state(v=(0.5, -0.4, 0.9), p=0.7)
    h = np.array([[0.9, 0.4, -1.0, 0.3, 0.2], [0.2, 0.05, 0.1, -0.6, 1.1]])
    out = _src(O, s, h)
    R = _rot(0.7, -0.4, 1.9)
    s2 = s.copy()
    s2[1:4] = R @ s[1:4]
    h2 = h.copy()
    h2[:, 2:5] = h[:, 2:5] @ R.T
    out2 = _src(O, s2, h2)
    np.testing.assert_allclose(out2[1:4], R @ out[1:4], rtol=0, atol=1e-13 * np.abs(out[1:4]).max())
    assert abs(out2[4] - out[4]) <= 1e-13 * abs(out[4])


def test_galilean_boost(O):
    s = _state(v=(0.5, -0.4, 0.9), p=0.7)
    h = np.array([[0.9, 0.4, -1.0, 0.3, 0.2], [0.2, 0.05, 0.1, -0.6, 1.1]])
    out = _src(O, s, h)
    W = np.array([0.8, -0.3, 0.45])
    s2 = s.copy()
    s2[1:4] += W
    h2 = h.copy()
    h2[:, 2:5] += W
    out2 = _src(O, s2, h2)
    scale = np.abs(out[1:4]).max()
    np.testing.assert_allclose(out2[1:4], out[1:4], rtol=0, atol=1e-13 * scale)
    assert abs(out2[4] - (out[4] + W @ out[1:4])) <= 1e-12 * max(abs(out[4]), scale)


# (rho, p, V_p, rho_H, p_H, V_H, tolerance): proton T_p = (1 + m_e/m_p) p / rho, T_H = 2 p_H / rho_H
MC_CASES = [
    (1.0, 4.0, (1.0, 0.0, 0.0), 1.0, 0.005, (0.0, 0.0, 0.0), 0.015),   # hot protons, cold atoms
    (1.0, 0.01, (1.0, 0.0, 0.0), 1.0, 2.0, (0.0, 0.0, 0.0), 0.015),    # cold protons, hot atoms
    (1.0, 1.0, (0.5, 0.2, 0.0), 1.0, 0.1, (0.0, -0.3, 0.1), 0.03),
    (1.0, 0.5, (0.3, 0.0, 0.0), 0.5, 0.25, (-1.0, 0.0, 0.0), 0.03),
    (0.2, 2.0, (3.0, 0.5, 0.0), 0.1, 0.02, (-0.5, 0.0, 0.0), 0.03),
]


@pytest.mark.parametrize("case", MC_CASES)
def test_bera_vs_maxwellian_monte_carlo(O, case):
    """With sigma = 1 (a_2 = 0) the Bera rates are closed-form approximations of Maxwellian
    averages: m = n_p n_H <g (w_H - w_p)>, E = n_p n_H <g (|w_H|^2 - |w_p|^2) / 2> over a proton
    Maxwellian of c^2 = T_p and an atom Maxwellian of c^2 = T_H.  They agree within 2%; a sign
    flip in the thermal drag terms T / U_m (V_p - V_H) or swapped temperatures inside U_m or U_E
    moves them by 3-25% in these cases."""
    rho, p, vp, rho_h, p_h, vh, tol = case
    vp, vh = np.array(vp), np.array(vh)
    s = _state(rho=rho, v=vp, p=p)
    out = O.bera_source(s[None], np.array([[[rho_h, p_h, *vh]]]), a2=0.0, n_h_lism=1.0, kn=1.0)[0]
    n_p = rho / (1.0 + ME_MP)
    Tp, Th = (1.0 + ME_MP) * p / rho, 2.0 * p_h / rho_h
    rng = np.random.default_rng(17)
    N = 2_000_000
    wp = rng.normal(0.0, np.sqrt(Tp / 2.0), (N, 3)) + vp
    wh = rng.normal(0.0, np.sqrt(Th / 2.0), (N, 3)) + vh
    g = np.linalg.norm(wp - wh, axis=1)
    m_mc = n_p * rho_h * (g[:, None] * (wh - wp)).mean(axis=0)
    e_mc = n_p * rho_h * (g * 0.5 * ((wh * wh).sum(1) - (wp * wp).sum(1))).mean()
    assert np.linalg.norm(out[1:4] - m_mc) <= tol * np.linalg.norm(m_mc), (out[1:4], m_mc)
    assert abs(out[4] - e_mc) <= tol * abs(e_mc), (out[4], e_mc)


# ---- the hydrogen fluids' outputs (Setka_physics.cpp:1005-1094, 1241-1250) --------------------
def _fluids(rng, n_h):
    return np.stack([np.concatenate([[rng.uniform(0.05, 0.3)], [rng.uniform(0.01, 0.1)], rng.uniform(-3, 1, 3)])
                     for _ in range(n_h)])


@pytest.mark.parametrize("zone_cell", [(0.0, 1), (0.0, 2), (100.0, 3), (100.0, 4)])
def test_fluid_sources_conserve_with_the_plasma(O, zone_cell):
    """Charge exchange swaps a proton and an atom: the atoms made equal the atoms lost (mass summed
    over the fluids is zero), and the fluids gain the momentum and energy the plasma loses; all
    hydrogen sources carry an extra 1/n_H_LISM."""
    rng = np.random.default_rng(11)
    q, ct = zone_cell
    s = _state(rho=0.9, v=(0.5, -0.1, 0.2), p=0.4)
    h = _fluids(rng, 4)
    out, hout = O.bera_source_h(s[None], h[None], scal=[(q * s[0], 0.1)], cell_type=[ct])
    out, hout = out[0], hout[0]
    ref = O.bera_source(s[None], h[None], scal=[(q * s[0], 0.1)], cell_type=[ct])[0]
    np.testing.assert_array_equal(out, ref)  # the plasma part is unchanged
    scale = np.abs(hout[:, 0]).max()
    assert abs(hout[:, 0].sum()) < 1e-14 * scale
    for i in range(1, 5):
        assert abs(out[i] + NH * hout[:, i].sum()) < 1e-13 * np.abs(out[i])
    assert np.abs(hout).min() >= 0.0 and np.abs(hout[:, 1:]).max() > 0.0


def test_fluid_sources_route_to_the_zone_fluid(O):
    """Without pickup ions the atoms made from every fluid join fluid `zone` (the hydrogen_arise
    tables of Phys_param.cpp:149-171): one fluid with atoms in zone 3 loses what fluid 3 gains."""
    s = _state()
    h = np.zeros((4, 5))
    h[:, 1] = 0.05
    h[1] = [0.2, 0.05, -2.0, 0.3, 0.1]
    out, hout = O.bera_source_h(s[None], h[None], scal=[(100.0 * s[0], 0.0)], cell_type=[3])
    hout = hout[0]
    assert np.all(hout[[0, 3]] == 0.0)
    assert hout[1, 0] < 0.0 and hout[2, 0] > 0.0
    assert abs(hout[1, 0] + hout[2, 0]) < 1e-15 * abs(hout[2, 0])
    # in zone 2 the same atoms join fluid 2, the one they came from: its mass source cancels
    out, h2 = O.bera_source_h(s[None], h[None], scal=[(0.0, 0.0)], cell_type=[2])
    assert abs(h2[0, 1, 0]) < 1e-15 * abs(hout[2, 0])
    # a caller's table sends zone 3's atoms to fluid 4
    ar = np.tile(np.arange(1, 5), (4, 1))
    ar[2] = 4
    out, h3 = O.bera_source_h(s[None], h[None], scal=[(100.0 * s[0], 0.0)], cell_type=[3], arise=ar)
    np.testing.assert_array_equal(h3[0, 3], hout[2])
    assert np.all(h3[0, 2] == 0.0)


def test_fluid_sources_cold_beam(O):
    """A cold beam crossing cold protons: fluid k loses sigma n_p n_H |dV| atoms per unit time with
    momentum n_H-weighted V_H, and the atoms made carry V_p (times n_H_LISM/Kn over n_H_LISM)."""
    rho, vp = 1.1, np.array([0.4, -0.3, 0.2])
    s = np.array([rho, *vp, 1e-12, 0.0, 0.0, 0.0])
    vh = np.array([-2.0, 0.5, 0.3])
    h = np.array([[0.3, 1e-12, *vh], [0.0, 0.0, 0.0, 0.0, 0.0]])
    out, hout = O.bera_source_h(s[None], h[None])   # zone 1: atoms made join fluid 1
    g = np.linalg.norm(vp - vh)
    sig = (1.0 - A2 * np.log(g)) ** 2
    n_p = rho / (1.0 + ME_MP)
    rate = sig * n_p * 0.3 * g * DDP / NH
    np.testing.assert_allclose(hout[0, 0, 0], 0.0, atol=1e-8 * rate)
    np.testing.assert_allclose(hout[0, 0, 1:4], rate * (vp - vh), rtol=1e-6)
