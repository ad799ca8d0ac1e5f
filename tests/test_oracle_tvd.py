"""Pins of the oracle's TVD path: the limiter (Help_funk.cpp:26-70), the neighbour choice
(Setka::Init_TVD, Setka_physics.cpp:648-744) and the second-order face states of
Setka::Snos_on_Gran (Setka_gran_snos.cpp:186-736).

A limited linear reconstruction must reproduce a linear field exactly, fall back to the cell value
at an extremum, and -- in the all-Zone_1 case, which interpolates r^2 rho, r^(2 gamma) p and the
spherical components of V and B -- carry a spherically symmetric wind onto the face exactly."""
import numpy as np
import pytest


def test_minmod_and_linear_known_answers(O):
    L = O.lib()
    assert L.or_minmod(2.0, 3.0) == 2.0
    assert L.or_minmod(-2.0, -0.5) == -0.5
    assert L.or_minmod(2.0, -3.0) == 0.0
    # |slope| below 1e-8 counts as sign 0: half the smaller magnitude survives
    assert L.or_minmod(5e-9, 4.0) == 0.5 * 5e-9
    # linear data: both one-sided slopes equal, the value at y is exact
    assert L.or_linear(-3.0, 1.0 - 0.6, -1.0, 1.0 - 0.2, 2.0, 1.0 + 0.4, 0.0) == pytest.approx(1.0, abs=1e-15)
    # extremum at x2: slopes of opposite sign, the cell value is returned
    assert L.or_linear(-3.0, 0.0, -1.0, 5.0, 2.0, 1.0, 0.0) == 5.0


def _cube(C, O, n=5, jitter=0.0):
    mesh, prim, params = C.cube_brio_wu(n=n, jitter=jitter)
    om = O.OracleMesh(mesh, tvd=True)
    g = O.measure(om, mesh.node_xyz)
    return mesh, params, om, g


def test_init_tvd_straight_lines_on_a_cube(C, O):
    mesh, params, om, g = _cube(C, O)
    tvd = om.arrs["face_tvd"]
    cc = g["cell_center"]
    n, fc = g["face_normal"], g["face_center"]
    h = 1.0 / 5
    checked = 0
    for f in range(mesh.n_faces):
        A, B = mesh.face_cells[f]
        if B < 0:
            assert tvd[f, 1] == -1
            continue
        behind_A = cc[A] - h * n[f]
        behind_B = cc[B] + h * n[f]
        ia = np.nonzero(np.all(np.abs(cc - behind_A) < 1e-9, axis=1))[0]
        ib = np.nonzero(np.all(np.abs(cc - behind_B) < 1e-9, axis=1))[0]
        if ia.size:
            assert tvd[f, 0] == ia[0]   # the cell straight behind A (cosine -1)
            checked += 1
        else:
            # A on the boundary: a side neighbour is picked (cosine -1/sqrt(5)) and kept (0.45 > 0.1)
            aa = tvd[f, 0]
            assert aa >= 0 and aa != B
            d = cc[aa] - fc[f]
            assert abs(np.dot(n[f], d / np.linalg.norm(d)) + 1 / np.sqrt(5)) < 1e-12
        if ib.size:
            assert tvd[f, 1] == ib[0]
    assert checked > 100


def _state_fn(x):
    s = np.zeros(x.shape[:-1] + (8,))
    s[..., 0] = 1.0 + 0.3 * x[..., 0] + 0.2 * x[..., 1]
    s[..., 1] = 0.5 - 0.4 * x[..., 2]
    s[..., 2] = 0.1 * x[..., 0]
    s[..., 3] = 0.2
    s[..., 4] = 2.0 + 0.5 * x[..., 1] - 0.3 * x[..., 2]
    s[..., 5] = 0.7 + 0.1 * x[..., 0]
    s[..., 6] = -0.4 * x[..., 1]
    s[..., 7] = 0.3
    return s


def test_linear_field_reconstructed_exactly(C, O):
    mesh, params, om, g = _cube(C, O)
    p = dict(params, tvd=1)
    prim = _state_fn(g["cell_center"])
    tvd = om.arrs["face_tvd"]
    cc, fc, n = g["cell_center"], g["face_center"], g["face_normal"]
    done = 0
    for f in range(mesh.n_faces):
        A, B = mesh.face_cells[f]
        if B < 0 or tvd[f, 0] < 0 or tvd[f, 1] < 0:
            continue
        # only faces whose four cells lie on one straight line (colinear: slopes are exact)
        line = np.array([cc[tvd[f, 0]], cc[A], cc[B], cc[tvd[f, 1]]]) - fc[f]
        if np.abs(np.cross(line, n[f])).max() > 1e-12:
            continue
        rc, qL, qR = O.face_states(om, p, prim, g, f)
        exp = _state_fn(fc[f])
        assert rc == 0
        np.testing.assert_allclose(qL, exp, rtol=0, atol=1e-14)
        np.testing.assert_allclose(qR, exp, rtol=0, atol=1e-14)
        done += 1
    assert done > 100
    # with TVD off the same faces keep the first-order cell values
    f = next(f for f in range(mesh.n_faces) if mesh.face_cells[f, 1] >= 0 and tvd[f, 0] >= 0 and tvd[f, 1] >= 0)
    rc, qL, qR = O.face_states(om, dict(params, tvd=0), prim, g, f)
    np.testing.assert_array_equal(qL, prim[mesh.face_cells[f, 0]])


def test_extremum_keeps_the_cell_value(C, O):
    mesh, params, om, g = _cube(C, O)
    tvd = om.arrs["face_tvd"]
    f = next(f for f in range(mesh.n_faces) if mesh.face_cells[f, 1] >= 0 and tvd[f, 0] >= 0 and tvd[f, 1] >= 0)
    A = mesh.face_cells[f, 0]
    prim = np.tile([1.0, 0.1, 0.2, 0.3, 1.0, 0.1, 0.2, 0.3], (mesh.n_cells, 1))
    prim[A, 0] = 3.0   # a density peak in A: the left state must not overshoot
    rc, qL, qR = O.face_states(om, dict(params, tvd=1), prim, g, f)
    assert qL[0] == 3.0


def test_positivity_falls_back_to_first_order(C, O):
    mesh, params, om, g = _cube(C, O)
    tvd = om.arrs["face_tvd"]
    cc, fc, n = g["cell_center"], g["face_center"], g["face_normal"]
    for f in range(mesh.n_faces):
        A, B = mesh.face_cells[f]
        if B >= 0 and tvd[f, 0] >= 0 and tvd[f, 1] >= 0:
            line = np.array([cc[tvd[f, 0]], cc[A], cc[B], cc[tvd[f, 1]]]) - fc[f]
            if np.abs(np.cross(line, n[f])).max() < 1e-12:
                break
    AA, A, B, BB = tvd[f, 0], mesh.face_cells[f, 0], mesh.face_cells[f, 1], tvd[f, 1]
    prim = np.tile([1.0, 0.1, 0.2, 0.3, 1.0, 0.1, 0.2, 0.3], (mesh.n_cells, 1))
    # a steep linear density profile falling through zero at the face
    d = np.array([np.dot(cc[k] - fc[f], n[f]) for k in (AA, A, B, BB)])
    for k, x in zip((AA, A, B, BB), d):
        prim[k, 0] = 1e-8 - 0.5 * x * 1e-3 / abs(d[1])
    rc, qL, qR = O.face_states(om, dict(params, tvd=1), prim, g, f)
    assert qL[0] == prim[A, 0] or qL[0] >= 1e-7
    assert qR[0] == prim[B, 0] or qR[0] >= 1e-7
    assert qL[0] == prim[A, 0] and qR[0] == prim[B, 0]


def test_zone1_wind_transported_by_tvd(C, O):
    mesh, _, params = C.shell(n_edge=4, n_layers=8, r_zone1_au=1e9)
    om = O.OracleMesh(mesh, tvd=True)
    g = O.measure(om, mesh.node_xyz)
    gamma = params["gamma"]
    from test_oracle_face_states import _wind
    tvd = om.arrs["face_tvd"]
    for axis, sel in ((2, True), (0, False)):
        prim = _wind(g["cell_center"], gamma, v_tan=0.4, b_tan=0.9, axis=axis)
        G = g["face_center"]
        theta = np.arccos(G[:, 2] / np.linalg.norm(G, axis=1))
        eq = (theta > np.pi / 9) & (theta < 8 * np.pi / 9)
        scale = np.abs(prim).max(axis=0)
        done = 0
        for f in np.nonzero((mesh.face_cells[:, 1] >= 0) & (tvd[:, 0] >= 0) & (tvd[:, 1] >= 0) & (eq == sel))[0]:
            cells = [tvd[f, 0], mesh.face_cells[f, 0], mesh.face_cells[f, 1], tvd[f, 1]]
            cc = g["cell_center"][cells]
            if np.min(np.linalg.norm(np.delete(cc, axis, axis=1), axis=1)) < 1e-6:
                continue
            rc, qL, qR = O.face_states(om, dict(params, tvd=1), prim, g, f)
            exp = _wind(G[f], gamma, v_tan=0.4, b_tan=0.9, axis=axis)
            for q, side in ((qL, cells[1]), (qR, cells[2])):
                e = exp.copy()
                for k in (0, 4):   # below 1e-7 the cell's own value is used (positivity)
                    if e[k] < 1e-7:
                        e[k] = prim[side, k]
                assert np.all(np.abs(q - e) <= 1e-12 * scale), (f, q, e)
            done += 1
        assert done > 20


@pytest.mark.parametrize("zones", [(2, 3), (3, 4)])
def test_zone_transition_cases_exact_on_linear_field(C, O, zones):
    """Across a Zone_2/3 (or 3/4) transition the reference switches to one-sided lines
    (Setka_gran_snos.cpp:649-713); each of them reproduces a linear field exactly."""
    mesh, _, params = C.cube_brio_wu(n=6, jitter=0.0)
    g0 = O.measure(O.OracleMesh(mesh), mesh.node_xyz)
    cx = g0["cell_center"][:, 0]
    mesh.cell_zone = np.where(cx < 0.45, zones[0], zones[1]).astype(np.int8)
    om = O.OracleMesh(mesh, tvd=True)
    g = O.measure(om, mesh.node_xyz)
    prim = _state_fn(g["cell_center"])
    tvd = om.arrs["face_tvd"]
    cc, fc, n = g["cell_center"], g["face_center"], g["face_normal"]
    seen = set()
    for f in range(mesh.n_faces):
        A, B = mesh.face_cells[f]
        if B < 0 or tvd[f, 0] < 0 or tvd[f, 1] < 0:
            continue
        line = np.array([cc[tvd[f, 0]], cc[A], cc[B], cc[tvd[f, 1]]]) - fc[f]
        if np.abs(np.cross(line, n[f])).max() > 1e-12:
            continue
        z = tuple(int(mesh.cell_zone[k]) for k in (tvd[f, 0], A, B, tvd[f, 1]))
        rc, qL, qR = O.face_states(om, dict(params, tvd=1), prim, g, f)
        exp = _state_fn(fc[f])
        np.testing.assert_allclose(qL, exp, rtol=0, atol=1e-14)
        np.testing.assert_allclose(qR, exp, rtol=0, atol=1e-14)
        seen.add(z)
    a, b = zones
    # the three transition patterns and the plain case on both sides all occurred
    for z in ((a, a, b, b), (a, a, a, b), (a, b, b, b), (a, a, a, a), (b, b, b, b)):
        assert z in seen, (z, seen)


def test_heliopause_flag_keeps_first_order(C, O):
    """A face flagged heliopause (face_flags bit2, Gran::type2 == HP) reconstructs to second order
    only with Snos_on_HP (Setka_gran_snos.cpp:188-189); the flag leaves the Riemann solver alone."""
    mesh, _, params = C.cube_brio_wu(n=5, jitter=0.1)
    g0 = O.measure(O.OracleMesh(mesh), mesh.node_xyz)
    prim = _state_fn(g0["cell_center"])
    prim[:, 0] += 0.3 * np.sin(7.0 * g0["cell_center"][:, 1])  # curvature: second order differs
    flags = np.zeros(mesh.n_faces, np.uint8)
    interior = np.nonzero(mesh.face_cells[:, 1] >= 0)[0]
    flags[interior[::3]] = 4
    mesh.face_flags = flags
    om = O.OracleMesh(mesh, tvd=True)
    g = O.measure(om, mesh.node_xyz)
    tvd = om.arrs["face_tvd"]
    n_diff = 0
    for f in interior[:60]:
        if tvd[f, 0] < 0 or tvd[f, 1] < 0:
            continue
        first = O.face_states(om, dict(params, tvd=0), prim, g, f)
        hp_off = O.face_states(om, dict(params, tvd=1, snos_on_hp=0), prim, g, f)
        hp_on = O.face_states(om, dict(params, tvd=1, snos_on_hp=1), prim, g, f)
        if flags[f]:
            np.testing.assert_array_equal(hp_off[1], first[1])
            np.testing.assert_array_equal(hp_off[2], first[2])
            n_diff += int(np.any(hp_on[1] != first[1]))
        else:
            np.testing.assert_array_equal(hp_off[1], hp_on[1])
            np.testing.assert_array_equal(hp_off[2], hp_on[2])
    assert n_diff > 3
    # the Riemann solver ignores bit2: same fluxes as unflagged faces with TVD off
    rc, fl4, _, _ = O.face_flux(om, dict(params, tvd=0), prim, g)
    mesh.face_flags = None
    rc, fl0, _, _ = O.face_flux(O.OracleMesh(mesh, tvd=True), dict(params, tvd=0), prim, g)
    np.testing.assert_array_equal(fl4, fl0)


def _banded_shell(C, O, n_edge=4, n_layers=12):
    """Shell whose radial layers are Zone_1 | Zone_2 | Zone_3 | Zone_4 in bands of three layers, so
    that a radial line AA - A -|- B - BB crosses every zone transition of the reconstruction."""
    mesh, _, params = C.shell(n_edge=n_edge, n_layers=n_layers, r_zone1_au=1e9)
    r_in = np.linalg.norm(mesh.node_xyz, axis=1)[mesh.cell_nodes].min(axis=1)  # inner sphere of the cell
    layer = np.searchsorted(np.unique(np.round(r_in, 9)), np.round(r_in, 9))
    mesh.cell_zone = (1 + np.minimum(layer // 3, 3)).astype(np.int8)
    om = O.OracleMesh(mesh, tvd=True)
    g = O.measure(om, mesh.node_xyz)
    return mesh, params, om, g


def _radial_faces(mesh, om, g):
    """Interior faces whose second neighbours lie on the face's radial line, with their zones."""
    tvd = om.arrs["face_tvd"]
    cc, fc = g["cell_center"], g["face_center"]
    out = []
    for f in range(mesh.n_faces):
        A, B = mesh.face_cells[f]
        if B < 0 or tvd[f, 0] < 0 or tvd[f, 1] < 0:
            continue
        cells = (tvd[f, 0], A, B, tvd[f, 1])
        u = fc[f] / np.linalg.norm(fc[f])
        pts = np.array([cc[k] for k in cells])
        if np.abs(np.cross(pts, u)).max() > 1e-9 * np.linalg.norm(pts, axis=1).max():
            continue
        out.append((f, cells, tuple(int(mesh.cell_zone[k]) for k in cells)))
    return out


TRANSITIONS = [(1, 1, 1, 2), (1, 1, 2, 2), (1, 2, 2, 2)]


def test_zone1_zone2_transitions_carry_the_wind(C, O):
    """Along a radial line a symmetric wind has constant V (radial 5 + azimuthal 0.4) in both
    spherical and Cartesian components, and r^(2 gamma) p is constant: wherever a transition case
    interpolates those (Setka_gran_snos.cpp:401-660) the face gets the wind's value exactly."""
    mesh, params, om, g = _banded_shell(C, O)
    from test_oracle_face_states import _wind
    gamma = params["gamma"]
    prim = _wind(g["cell_center"], gamma, v_tan=0.4, b_tan=0.0, axis=2)
    seen = set()
    for f, cells, z in _radial_faces(mesh, om, g):
        if z not in TRANSITIONS:
            continue
        # keep away from the polar axis, where the azimuthal direction is undefined
        if np.linalg.norm(g["face_center"][f][:2]) < 1e-3 * np.linalg.norm(g["face_center"][f]):
            continue
        rc, qL, qR = O.face_states(om, dict(params, tvd=1), prim, g, f)
        exp = _wind(g["face_center"][f], gamma, v_tan=0.4, b_tan=0.0, axis=2)
        tol = 1e-12 * np.abs(prim).max(axis=0)
        # V: spherical lines on the Zone_1 side, Cartesian lines on the plain side -- both exact
        assert np.all(np.abs(qL[1:4] - exp[1:4]) <= tol[1:4]), (z, qL, exp)
        assert np.all(np.abs(qR[1:4] - exp[1:4]) <= tol[1:4]), (z, qR, exp)
        # p: scaled by r^(2 gamma) on the spherical sides (left of 1112 and 1122, right of 1112)
        if z != (1, 2, 2, 2):
            assert abs(qL[4] - exp[4]) <= tol[4], (z, qL[4], exp[4])
        if z == (1, 1, 1, 2):
            assert abs(qR[4] - exp[4]) <= tol[4], (z, qR[4], exp[4])
        seen.add(z)
    assert seen == set(TRANSITIONS), seen


def test_zone1_zone2_transitions_line_choice(C, O):
    """Zig-zag densities 1, 3, 2, 4 on AA, A, B, BB (an extremum in A and in B) tell the lines
    apart: a minmod-limited line through three cells returns the middle cell's value, the line
    through two cells interpolates or extrapolates linearly in the distance.  Density is
    interpolated unscaled in all three transition cases."""
    mesh, params, om, g = _banded_shell(C, O)
    from test_oracle_face_states import _wind
    base = _wind(g["cell_center"], params["gamma"], v_tan=0.0, b_tan=0.0, axis=2)
    cc, fc = g["cell_center"], g["face_center"]
    rho = (1.0, 3.0, 2.0, 4.0)
    seen = set()
    for f, cells, z in _radial_faces(mesh, om, g):
        if z not in TRANSITIONS:
            continue
        prim = base.copy()
        for k, v in zip(cells, rho):
            prim[k, 0] = v
        AA, A, B, BB = cells
        d1 = np.linalg.norm(cc[A] - fc[f])
        d2 = np.linalg.norm(cc[B] - fc[f])
        dd1 = np.linalg.norm(cc[A] - cc[AA])
        dd2 = np.linalg.norm(cc[B] - cc[BB])
        limited_left = rho[1]                                          # extremum at A
        limited_right = rho[2]                                         # extremum at B
        interp = rho[1] + (rho[2] - rho[1]) * d1 / (d1 + d2)           # A-B line at the face
        extrap_left = rho[1] + (rho[1] - rho[0]) * d1 / dd1            # AA-A line
        extrap_right = rho[2] + (rho[2] - rho[3]) * d2 / dd2           # B-BB line
        expect = {(1, 1, 1, 2): (limited_left, interp),
                  (1, 1, 2, 2): (extrap_left, extrap_right),
                  (1, 2, 2, 2): (interp, limited_right)}[z]
        rc, qL, qR = O.face_states(om, dict(params, tvd=1), prim, g, f)
        assert qL[0] == pytest.approx(expect[0], rel=1e-13), (z, qL[0], expect)
        assert qR[0] == pytest.approx(expect[1], rel=1e-13), (z, qR[0], expect)
        seen.add(z)
    assert seen == set(TRANSITIONS), seen
