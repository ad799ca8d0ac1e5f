"""The oracle's HP smoothing on the ray grids and the corrections after the move (Setka_move.cpp:
700-1357), pinned by what fixes them: points on a circle give the circle fit's centre
(fitCircleRobust, :2544-2620) and no angle pull; points on a parabola (numpy.polyfit, not the
reference's closed form) give no macros1 pull and set the size of every other one; the tail and BS
corrections leave the corrected nodes on the line or radius they name; and the set-up order of
TITAN.cpp:108-112 (Smooth_angle_HP adds to the sets Smooth_head_HP3 later clears near the nose)."""
import math

import numpy as np
import pytest

from helpers import hp_grid_lists


def _lm_circle(P, iters=8, tol=1e-8):
    """fitCircleRobust restated in numpy: LM from the origin; the damping grows tenfold every step."""
    p = np.array([0.0, 0.0, np.mean(np.hypot(P[:, 0], P[:, 1]))])
    lam, prev = 0.001, np.finfo(float).max
    for _ in range(iters):
        d = P - p[:2]
        dist = np.hypot(d[:, 0], d[:, 1])
        res = dist - p[2]
        J = np.stack([-d[:, 0] / dist, -d[:, 1] / dist, -np.ones(len(P))], axis=1)
        err = math.sqrt((res * res).sum() / len(P))
        if abs(prev - err) < tol:
            break
        prev = err
        H = J.T @ J
        H[np.diag_indices(3)] *= 1.0 + lam
        p = p + np.linalg.solve(H, -J.T @ res)
        lam *= 10.0
    return p[:2], p[2]


def test_circle_fit_exact_on_circle(O):
    """Seven points around a whole circle: the fit stops (rms change below 1e-8) within 5e-8 R of
    the centre, the only output Smooth_angle_HP keeps; the radius, whose steps the growing damping
    shrinks most, within 1e-5 R."""
    rng = np.random.default_rng(1)
    for _ in range(10):
        c = rng.uniform(-2, 2, 2)
        R = rng.uniform(20, 40)
        a = rng.uniform(0, 2 * math.pi) + 2 * math.pi / 7 * np.arange(7)
        P = c + R * np.stack([np.cos(a), np.sin(a)], axis=1)
        cc, rr = O.circle_fit(P)
        np.testing.assert_allclose(cc, c, rtol=0, atol=5e-8 * R)
        assert abs(rr - R) < 1e-5 * R


def test_circle_fit_follows_the_damping_rule(O):
    """A short arc far from the origin: 8 damped steps stop short of the circle, where a plain
    Gauss-Newton would reach it; the oracle stops where the numpy restatement of the rule does."""
    c, R = np.array([25.0, -10.0]), 12.0
    a = 0.4 + 0.05 * np.arange(7)
    P = c + R * np.stack([np.cos(a), np.sin(a)], axis=1)
    P += np.random.default_rng(2).normal(0, 1e-3, P.shape)
    cc, rr = O.circle_fit(P)
    cn, rn = _lm_circle(P)
    np.testing.assert_allclose(cc, cn, rtol=0, atol=1e-9 * R)
    assert np.abs(cc - c).max() > 1e-3


def _pts_on_parabola(coef, xs, phis):
    rho = np.polyval(coef, xs)
    return np.stack([xs, rho * np.cos(phis), rho * np.sin(phis)], axis=1)


def _node_set(rng):
    """A node and its 8 grid neighbours and 6 ring neighbours in general position."""
    A = np.array([3.0, 20.0 * math.cos(0.7), 20.0 * math.sin(0.7)])
    xs = A[0] + np.array([-4.0, -8.5, 4.2, 8.9])
    along = _pts_on_parabola([0.01, -0.2, 21.0], xs, 0.7 + rng.uniform(-0.02, 0.02, 4))
    along[:, 1:] *= (1.0 + rng.uniform(-0.05, 0.05, 4))[:, None]
    ts = 0.7 + np.array([-0.21, -0.40, 0.19, 0.42])
    rho = 20.0 * (1.0 + rng.uniform(-0.05, 0.05, 4))
    az = np.stack([A[0] + rng.uniform(-1, 1, 4), rho * np.cos(ts), rho * np.sin(ts)], axis=1)
    P8 = np.concatenate([along, az])  # AA1, AA11, AA3, AA33, AA2, AA22, AA4, AA44
    ring = np.stack([A[0] + rng.uniform(-1, 1, 6), 19.0 * np.cos(0.7 + 0.2 * np.arange(1, 7)),
                     19.0 * np.sin(0.7 + 0.2 * np.arange(1, 7))], axis=1)
    return A, P8, ring


def _rho(p):
    return math.hypot(p[1], p[2])


def test_macros1_and_angle_terms(O):
    rng = np.random.default_rng(3)
    A, P8, ring = _node_set(rng)
    cy, cz, dt, ca, cl, cg = 0.3, -0.2, 0.01, 1e-4, 0.1 * 1.0 * 0.03, 0.1 * 1.8 * 0.03
    rc, V = O.hp_terms(A, P8, (cy, cz), ring, ca, cl, cg, dt)
    assert rc == 0
    # macros1: the parabola through three (abscissa, r_yz) points by polyfit, evaluated at the node
    ang = lambda p: math.atan2(p[2], p[1]) % (2 * math.pi)
    sets = [([P8[2], P8[0], P8[1]], lambda p: p[0], cl), ([P8[3], P8[2], P8[0]], lambda p: p[0], cl),
            ([P8[4], P8[6], P8[7]], ang, cg), ([P8[6], P8[4], P8[5]], ang, cg)]
    for q, (pts, absc, c) in enumerate(sets):
        coef = np.polyfit([absc(p) for p in pts], [_rho(p) for p in pts], 2)
        r = np.polyval(coef, absc(A))
        want = c * (A * np.array([0.0, r / _rho(A) - 1.0, r / _rho(A) - 1.0])) / dt
        np.testing.assert_allclose(V[1 + q], want, rtol=1e-9, atol=1e-12)
    # the angle term: towards the circle about the centre through the ring's mean distance
    rm = np.mean([math.hypot(p[1] - cy, p[2] - cz) for p in ring])
    d = A[1:] - [cy, cz]
    tgt = np.array([cy, cz]) + rm * d / np.linalg.norm(d)
    np.testing.assert_allclose(V[0], [0.0, *(ca * (tgt - A[1:]) / dt)], rtol=1e-12, atol=1e-15)


def test_no_pull_on_a_parabola_and_a_circle(O):
    rng = np.random.default_rng(4)
    A, P8, ring = _node_set(rng)
    coef = [0.01, -0.2, 21.0]
    P8[:4] = _pts_on_parabola(coef, P8[:4, 0], rng.uniform(0, 6, 4))
    rA = np.polyval(coef, A[0])
    # the node sits on both parabolas: the angular one takes rA at the node's azimuth t0
    t0 = 0.9
    tc = [0.5, -0.3, rA - 0.5 * t0 * t0 + 0.3 * t0]
    ts = t0 + np.array([-0.21, -0.40, 0.19, 0.42])  # AA2, AA22, AA4, AA44
    rho_t = np.polyval(tc, ts)
    P8[4:] = np.stack([rng.uniform(-1, 1, 4), rho_t * np.cos(ts), rho_t * np.sin(ts)], axis=1)
    A = np.array([A[0], rA * math.cos(t0), rA * math.sin(t0)])
    R = 17.0
    ring = np.stack([rng.uniform(-1, 1, 6), R * np.cos(t0 + 0.1 * np.arange(1, 7)),
                     R * np.sin(t0 + 0.1 * np.arange(1, 7))], axis=1)
    A2 = A.copy()
    A2[1:] = R * np.array([math.cos(t0), math.sin(t0)])
    rc, V = O.hp_terms(A, P8, (0.0, 0.0), ring, 1e-3, 0.003, 0.0054, 0.01)
    assert rc == 0
    assert np.abs(V[1:]).max() < 1e-10
    rc, V2 = O.hp_terms(A2, P8, (0.0, 0.0), ring, 1e-3, 0.003, 0.0054, 0.01)
    assert np.abs(V2[0]).max() < 1e-12


def test_repeated_abscissa_is_an_error(O):
    A, P8, ring = _node_set(np.random.default_rng(5))
    P8[1] = P8[0]  # AA11 = AA1
    rc, _ = O.hp_terms(A, P8, (0.0, 0.0), ring, 1e-3, 0.003, 0.0054, 0.01)
    assert rc == -1


def _shell(C, O):
    mesh, prim, params, rays = C.shell_fitted(zone1=False, noise=0.01)
    F = O.OracleFit(mesh, rays)
    return mesh, prim, params, rays, F


def test_corrections_after_the_move(C, O):
    mesh, prim, params, rays, F = _shell(C, O)
    x0 = mesh.node_xyz
    L = hp_grid_lists(rays, x0)
    N4 = 6
    F.set_hp_grid([], None, tail=L["tail"], N4=N4, bs_ref=L["bs_ref"], bs_lists=L["bs_lists"])
    om = O.OracleMesh(mesh)
    p = dict(params, sglag_TS=0, sglag_HP=0)
    rc, xn, _ = O.culc_velocity_surface(om, p, F, x0, prim, 1e-3)
    assert rc == 0
    rho = lambda v: math.hypot(xn[v, 1], xn[v, 2])
    for t in L["tail"]:
        h1, h2 = rho(t[N4 - 4]), rho(t[N4 - 6])
        x1, x2 = xn[t[N4 - 4], 0], xn[t[N4 - 6], 0]
        want = h2 + (h1 - h2) / (x1 - x2) * (xn[t[N4 - 5], 0] - x2)
        assert abs(rho(t[N4 - 5]) - want) < 1e-12 * want
        for k in range(N4 - 3, len(t)):
            assert abs(rho(t[k]) - h1) < 1e-13 * h1
    for ref, lst in zip(L["bs_ref"], L["bs_lists"]):
        for v in lst:
            assert abs(rho(v) - rho(ref)) < 1e-13 * rho(ref)
    # without move_HP / move_BS neither correction runs
    rc, xs, _ = O.culc_velocity_surface(om, dict(p, move_HP=0, move_BS=0), F, x0, prim, 1e-3)
    moved = np.concatenate([L["tail"].reshape(-1), *[np.asarray(b) for b in L["bs_lists"]]])
    np.testing.assert_array_equal(xs[moved], x0[moved])


def test_angle_sets_follow_the_setup_order(C, O):
    """TITAN.cpp:108-112: Smooth_angle_HP, then Smooth_head_HP3 clears and refills the sets of the
    nodes near the nose; run the other way round, those sets keep both."""
    mesh, prim, params, rays = C.shell_fitted(zone1=False, noise=0.01)
    rays = dict(rays, phi_a=1.0)  # the sphere smoothing up to 1 rad + 1 degree from +x
    F = O.OracleFit(mesh, rays)
    x0 = mesh.node_xyz
    L = hp_grid_lists(rays, x0, the_lo=0.0)
    F.smooth_angle_hp(L["angle_nodes"], L["nb6"], x0)
    F.smooth(2, x0)
    G = O.OracleFit(mesh, rays)
    G.smooth(2, x0)
    H = O.OracleFit(mesh, rays)
    H.smooth(2, x0)
    H.smooth_angle_hp(L["angle_nodes"], L["nb6"], x0)
    nose = [i for i, v in enumerate(F.sn) if F.sn_kind[i] == 2 and G.nb[i]]
    far = [F._sn_index[int(v)] for v in L["angle_nodes"] if not G.nb[F._sn_index[int(v)]]]
    assert nose and far
    for i in nose:
        assert F.nb[i] == G.nb[i]
    for k, v in enumerate(L["angle_nodes"]):
        i = F._sn_index[int(v)]
        if i in nose:
            assert set(H.nb[i]) == set(G.nb[i]) | set(L["nb6"][k].tolist())
        else:
            assert F.nb[i] == sorted(set(L["nb6"][k].tolist()))
