"""The oracle's B, C, D, E, H and G ray laws (Luch::dvigenie, Luch.cpp:153-739) against what fixes
them independently of their coefficient formulas: the Hermite end conditions of every bent
segment (textbook Hermite basis), the end values and slopes that define each spacing law, the
straight inner runs of B and C rays, the planes x = const of D and E rays, the C tail's uniform
run to L7, the int truncation of dd5 and dd6 (Luch.cpp:166-167), which nodes and coordinates each
type leaves alone, and the family order (Setka.cpp:1850-1880): H and G rays bridge nodes that B
and D rays move in the same call."""
import math

import numpy as np
import pytest

from helpers import ext_ray_set

ATOL = 1e-9  # coordinates up to 600: a relative 2e-12


# textbook cubic Hermite basis on [0, 1]
def _herm(u, p0, t0, p1, t1):
    h00, h10 = 2 * u ** 3 - 3 * u ** 2 + 1, u ** 3 - 2 * u ** 2 + u
    h01, h11 = -2 * u ** 3 + 3 * u ** 2, u ** 3 - u ** 2
    return h00 * np.asarray(p0) + h10 * np.asarray(t0) + h01 * np.asarray(p1) + h11 * np.asarray(t1)


def _sp1(s, a, b):
    """The cubic through (0, 0) and (1, 1) with slopes a and b there."""
    return float(_herm(s, [0.0], [a], [1.0], [b])[0])


def _sp2(s, a, b):
    """The quartic with p(0) = 0, p'(0) = a, p''(0) = 0, p(1) = 1, p'(1) = b."""
    A = np.array([[0, 0, 0, 0, 1], [0, 0, 0, 1, 0], [0, 0, 2, 0, 0], [1, 1, 1, 1, 1], [4, 3, 2, 1, 0]], float)
    c = np.linalg.solve(A, [0.0, a, 0.0, 1.0, b])
    return float(np.polyval(c, s))


def _pol(r, the, phi):
    return r * np.array([math.cos(the), math.sin(the) * math.cos(phi), math.sin(the) * math.sin(phi)])


def _cyl(x, rho, phi):
    return np.array([x, rho * math.cos(phi), rho * math.sin(phi)])


def _rho(p):
    return math.hypot(p[1], p[2])


def _close(got, want):
    np.testing.assert_allclose(got, want, rtol=0, atol=ATOL)


def check_ext(types, lists, opor, ang, g, x0, x1):
    """Every law, read off the moved nodes x1 (start x0)."""
    M0, M1, M11, M2, M3, M4 = (g[k] for k in ("M0", "M1", "M11", "M2", "M3", "M4"))
    R0, R1, R5 = g["R0"], g["R1"], g["R5"]
    written = np.zeros(len(x0), bool)
    for k, t in enumerate(types):
        N, O = lists[k], opor[k]
        the, phi = ang[k]
        if t in (1, 2):
            assert np.array_equal(x1[O[1]], x0[O[1]])  # the TS node is the fit's
            R2 = float(np.linalg.norm(x1[O[1]]))
            for j in range(M0):
                _close(x1[N[j]], _pol(R0 + j * (R1 - R0) / M0, the, phi))
            dr = (R2 - R1) / (M1 + 1)
            for j in range(M1):
                _close(x1[N[M0 + 1 + j]], _pol(R1 + (j + 1) * dr, the, phi))
            for j in range(M11):
                _close(x1[N[M0 + M1 + 2 + j]], _pol(R2 + (j + 1) * dr, the, phi))
            rl = R2 + M11 * dr
            p0 = [rl * math.cos(the), rl * math.sin(the)]
            base = M0 + M1 + M11 + 2
            w = list(range(M0)) + list(range(M0 + 1, M0 + M1 + 1)) + list(range(M0 + M1 + 2, base))
            if t == 1:
                _close(x1[N[M0]], _pol(R1, the, phi))
                w.append(M0)
                x3 = p0[0]
                for q in (2, 3):  # x set to x3, y and z kept
                    _close(x1[O[q]][0], x3)
                    assert np.array_equal(x1[O[q]][1:], x0[O[q]][1:])
                R3, R4 = _rho(x1[O[2]]), _rho(x1[O[3]])
                n = M2 - M11
                for j in range(n):
                    u = _sp1((j + 1) / (n + 1), g["ba1"], g["ba2"])
                    c = _herm(u, p0, [g["dd3"] * math.cos(the), g["dd3"] * math.sin(the)], [x3, R3], [0.0, g["dd4"]])
                    _close(x1[N[base + j]], _cyl(c[0], c[1], phi))
                    w.append(base + j)
                iHP, iBS = M0 + M1 + M2 + 2, M0 + M1 + M2 + M3 + 3
                for j in range(M3):
                    _close(x1[N[iHP + 1 + j]], _cyl(x3, R3 + _sp1((j + 1) / (M3 + 1), g["ba3"], g["ba4"]) * (R4 - R3),
                                                    phi))
                    w.append(iHP + 1 + j)
                for j in range(M4):
                    _close(x1[N[iBS + 1 + j]], _cyl(x3, R4 + (R5 - R4) * _sp2((j + 1) / (M4 + 1), g["ba5"], 2.0),
                                                    phi))
                    w.append(iBS + 1 + j)
                _close(x1[N[iBS + M4 + 1]], _cyl(x3, R5, phi))
                w.append(iBS + M4 + 1)
            else:
                d5, d6 = int(g["dd5"]), int(g["dd6"])
                N4, N5 = g["N4"], g["N5"]
                last = None
                for j in range(N4):
                    c = _herm((j + 1) / (N4 + 1), p0, [(math.cos(the) - 1) / 2 * d5, math.sin(the) * d5],
                              [g["L6"], p0[1]], [-float(d6), 0.0])
                    _close(x1[N[base + j]], _cyl(c[0], c[1], phi))
                    w.append(base + j)
                    last = x1[N[base + j]]
                for j in range(N5):
                    p = x1[N[base + N4 + j]]
                    _close(p[0], g["L6"] + j * (g["L7"] - g["L6"]) / (N5 - 1))
                    assert np.array_equal(p[1:], last[1:])
                    w.append(base + N4 + j)
                assert np.array_equal(x1[N[M0]], x0[N[M0]])  # the C ray leaves its R1 node
        elif t == 3:
            xD = x1[O[0]][0]
            for q in (1, 2):
                _close(x1[O[q]][0], xD)
                assert np.array_equal(x1[O[q]][1:], x0[O[q]][1:])
            y0, y1, y2 = _rho(x1[O[0]]), _rho(x1[O[1]]), _rho(x1[O[2]])
            n = M2 - M11
            w = []
            for j in range(n):
                _close(x1[N[1 + j]], _cyl(xD, y0 + (y1 - y0) * _sp1((j + 1) / (n + 1), 0.1, 0.3), phi))
                w.append(1 + j)
            for j in range(M3):
                _close(x1[N[n + 2 + j]], _cyl(xD, y1 + (y2 - y1) * _sp1((j + 1) / (M3 + 1), g["md1"], g["md2"]), phi))
                w.append(n + 2 + j)
            for j in range(M4):
                _close(x1[N[n + M3 + 3 + j]], _cyl(xD, y2 + (R5 - y2) * _sp2((j + 1) / (M4 + 1), g["md3"], 2.0), phi))
                w.append(n + M3 + 3 + j)
            _close(x1[N[n + M3 + M4 + 3]], _cyl(xD, R5, phi))
            w.append(n + M3 + M4 + 3)
            assert np.array_equal(x1[N[0]][1:], x0[N[0]][1:]) or N[0] in [v for L in lists for v in L[1:]]
        elif t == 4:
            xe = x1[O[0]][0]
            for i in (0, M3 + 1):
                _close(x1[N[i]][0], xe)
                assert np.array_equal(x1[N[i]][1:], x0[N[i]][1:])
            y0, y1 = _rho(x1[O[1]]), _rho(x1[O[2]])
            w = [0, M3 + 1]
            for j in range(M3):
                _close(x1[N[1 + j]], _cyl(xe, y0 + (y1 - y0) * _sp1((j + 1) / (M3 + 1), g["ea1"], g["ea2"]), phi))
                w.append(1 + j)
            for j in range(M4):
                _close(x1[N[M3 + 2 + j]], _cyl(xe, y1 + (R5 - y1) * _sp2((j + 1) / (M4 + 1), g["ea3"], 2.0), phi))
                w.append(M3 + 2 + j)
            _close(x1[N[M3 + M4 + 2]], _cyl(xe, R5, phi))
            w.append(M3 + M4 + 2)
        else:
            z = [np.array([x1[v][0], _rho(x1[v])]) for v in O[:4] if v >= 0]
            if t == 5:
                z3, z1, z4, z2 = z
                p0, p1, t0, t1 = z1, z2, g["dd7"] * (z1 - z3), g["dd8"] * (z4 - z2)
                n = 4
            else:
                z3, z1, z2 = z
                p0, p1, t0, t1 = z1, z2, g["dd1"] * (z1 - z3), np.array([0.0, g["dd2"]])
                n = M2 - M11 - g["MF"] + 1
            w = []
            for j in range(n):
                c = _herm((j + 1) / (n + 1), p0, t0, p1, t1)
                _close(x1[N[1 + j]], _cyl(c[0], c[1], phi))
                w.append(1 + j)
        for i in w:
            written[N[i]] = True
        if t == 1:
            written[O[2]] = written[O[3]] = True
        if t == 3:
            written[O[1]] = written[O[2]] = True
    assert np.array_equal(x1[~written], x0[~written])


def _case(O, seed, counts=(6, 6, 6, 6, 6, 6), serial=(), geo=None, n_nodes=2000):
    rng = np.random.default_rng(seed)
    types, lists, opor, ang, g, xyz = ext_ray_set(rng, 0, n_nodes, counts, serial, geo)
    ext = O.RaysExt(types, lists, opor, ang, g)
    return types, lists, opor, ang, g, xyz, O.dvigenie_ext(ext, xyz)


@pytest.mark.parametrize("seed", [1, 2])
def test_every_family_follows_its_law(O, seed):
    types, lists, opor, ang, g, x0, x1 = _case(O, seed)
    check_ext(types, lists, opor, ang, g, x0, x1)


def test_rays_that_read_each_others_nodes_run_in_order(O):
    types, lists, opor, ang, g, x0, x1 = _case(O, 3, serial=(5, 6))
    check_ext(types, lists, opor, ang, g, x0, x1)


def test_c_rays_take_dd5_dd6_as_integers(O):
    """Luch.cpp:166-167 assigns dd5, dd6 to int: 19.7 and 14.9 act as 19 and 14."""
    a = _case(O, 4, geo=dict(dd5=19.7, dd6=14.9))
    b = _case(O, 4, geo=dict(dd5=19.0, dd6=14.0))
    np.testing.assert_array_equal(a[-1], b[-1])
    check_ext(*a)
    c = _case(O, 4, geo=dict(dd5=20.0, dd6=14.0))
    assert np.abs(c[-1] - b[-1]).max() > 1e-3


def test_laws_are_not_trivial(O):
    """The checks bite: the tangent scale changes the bent nodes, the spacing slopes the runs."""
    base = _case(O, 5)
    for key in ("dd3", "dd4", "dd7", "dd8", "dd1", "dd2", "ba1", "ba3", "ba5", "md1", "md3", "ea1", "ea3"):
        other = _case(O, 5, geo={key: base[4][key] * 1.25 + 0.1})
        assert np.abs(other[-1] - base[-1]).max() > 1e-6, key
