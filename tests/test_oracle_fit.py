"""The oracle's fitted-surface motion (or_culc_velocity_surface, or_dvigenie; Setka_move.cpp:111-1357,
Luch.cpp:6-152) against what geometry and the Riemann fan fix without the oracle's formulas:
straight rays, evenly spaced nodes where the ray law is linear (da = 1), the end slopes of its
polynomials, the wave speeds of a uniform state (S_L = v_n - a, S_M = v_n, S_R = v_n + a for B = 0),
the sphere and neighbour-mean smoothing as geometric projections, the cylindrical move of the
heliopause behind the Sun, the surface-neighbour count of a partial surface, and free-stream
preservation of the whole moving step."""
import math

import numpy as np
import pytest

GAMMA = 5.0 / 3.0


def _uniform(mesh, rho=1.3, p=0.7, v=(0.0, 0.0, 0.0)):
    prim = np.zeros((mesh.n_cells, 8))
    prim[:, 0] = rho
    prim[:, 1:4] = v
    prim[:, 4] = p
    return prim


def _fit(O, mesh, rays, x0=None, kinds=(1, 2, 3)):
    F = O.OracleFit(mesh, rays)
    for k in kinds:
        F.smooth(k, mesh.node_xyz if x0 is None else x0)
    return F


def _ray_radii(x, rays):
    return np.linalg.norm(x[rays["ray_nodes"]], axis=2)


def _perturb_surfaces(mesh, rays, seed=3, amp=0.05):
    """Moves every ray's TS, HP and BS node radially by a random factor."""
    rng = np.random.default_rng(seed)
    x = mesh.node_xyz.copy()
    M0, M1, M2, M3 = rays["M0"], rays["M1"], rays["M2"], rays["M3"]
    for j in (M0 + M1 + 1, M0 + M1 + M2 + 2, M0 + M1 + M2 + M3 + 3):
        ids = rays["ray_nodes"][:, j]
        x[ids] *= (1.0 + amp * rng.uniform(-1, 1, len(ids)))[:, None]
    return x


def test_rays_straight_and_even_where_linear(C, O):
    """da1..da4 = 1 make the TS->HP and HP->BS laws linear (ss(s) = s), so the nodes are evenly
    spaced between the moved surfaces; every node stays on its ray's direction."""
    mesh, prim, params, rays = C.shell_fitted(n_edge=4, M=(3, 5, 2, 7, 6, 4), da=(1.0, 1.0, 1.0, 1.0, 0.5))
    F = _fit(O, mesh, rays, kinds=())
    x = _perturb_surfaces(mesh, rays)
    y = O.dvigenie(F, x)
    R = _ray_radii(y, rays)
    M0, M1, M11, M2, M3, M4 = (rays[k] for k in ("M0", "M1", "M11", "M2", "M3", "M4"))
    iTS, iHP, iBS = M0 + M1 + 1, M0 + M1 + M2 + 2, M0 + M1 + M2 + M3 + 3
    # surfaces and the R1 node untouched
    for j in (M0, iTS, iHP, iBS):
        ids = rays["ray_nodes"][:, j]
        np.testing.assert_array_equal(y[ids], x[ids])
    # direction of every other node = (cos the, sin the cos phi, sin the sin phi)
    the, phi = rays["ray_angles"][:, 0], rays["ray_angles"][:, 1]
    u = np.stack([np.cos(the), np.sin(the) * np.cos(phi), np.sin(the) * np.sin(phi)], axis=1)
    for j in range(R.shape[1]):
        if j in (M0, iTS, iHP, iBS):
            continue
        ids = rays["ray_nodes"][:, j]
        np.testing.assert_allclose(y[ids], R[:, j:j + 1] * u, rtol=0, atol=1e-12 * R.max())
    R2, R3, R4 = R[:, iTS], R[:, iHP], R[:, iBS]
    dr = (R2 - rays["R1"]) / (M1 + 1)
    # R1 -> TS: M1 + 1 equal steps; M11 more of the same step; then linear to the HP; linear to the BS
    seg = R[:, M0:iTS + 1]
    np.testing.assert_allclose(np.diff(seg, axis=1), dr[:, None] * np.ones((1, M1 + 1)), rtol=1e-12)
    np.testing.assert_allclose(np.diff(R[:, iTS:iTS + M11 + 1], axis=1), dr[:, None] * np.ones((1, M11)), rtol=1e-12)
    start = R2 + dr * M11
    steps = (R3 - start) / (M2 - M11 + 1)
    np.testing.assert_allclose(np.diff(R[:, iTS + M11:iHP + 1], axis=1), steps[:, None] * np.ones((1, M2 - M11 + 1)),
                               rtol=1e-10)
    np.testing.assert_allclose(np.diff(R[:, iHP:iBS + 1], axis=1), ((R4 - R3) / (M3 + 1))[:, None] * np.ones((1, M3 + 1)),
                               rtol=1e-10)
    # R0 -> R1: M0 equal steps; last node at R5
    np.testing.assert_allclose(R[:, :M0 + 1], np.linspace(rays["R0"], rays["R1"], M0 + 1)[None, :] * np.ones((len(R2), 1)),
                               rtol=1e-13)
    np.testing.assert_allclose(R[:, -1], rays["R5"], rtol=1e-14)


@pytest.mark.parametrize("a,b", [(2.0, 0.5), (0.6, 1.7)])
def test_ray_law_end_slopes(C, O, a, b):
    """With many nodes the TS->HP law's end slopes approach da1 and da2, the HP->BS law's da3 and
    da4, and the BS->R5 law's da5 and 2 (the derivatives of the polynomials at s = 0 and 1)."""
    M = (2, 3, 1, 401, 301, 301)
    mesh, prim, params, rays = C.shell_fitted(n_edge=1, M=M, da=(a, b, a, b, a))
    F = _fit(O, mesh, rays, kinds=())
    R = _ray_radii(O.dvigenie(F, mesh.node_xyz), rays)[0]
    M0, M1, M11, M2, M3, M4 = M
    iTS, iHP, iBS = M0 + M1 + 1, M0 + M1 + M2 + 2, M0 + M1 + M2 + M3 + 3
    dr = (R[iTS] - rays["R1"]) / (M1 + 1)
    L4, n4 = R[iHP] - (R[iTS] + dr * M11), M2 - M11 + 1
    s0 = (R[iTS + M11 + 1] - R[iTS + M11]) / (L4 / n4)
    s1 = (R[iHP] - R[iHP - 1]) / (L4 / n4)
    assert abs(s0 - a) < 0.02 * a and abs(s1 - b) < 0.02 * max(b, 1)
    L5, n5 = R[iBS] - R[iHP], M3 + 1
    assert abs((R[iHP + 1] - R[iHP]) / (L5 / n5) - a) < 0.02 * a
    assert abs((R[iBS] - R[iBS - 1]) / (L5 / n5) - b) < 0.02 * max(b, 1)
    L6, n6 = rays["R5"] - R[iBS], M4 + 1
    assert abs((R[iBS + 1] - R[iBS]) / (L6 / n6) - a) < 0.02 * a
    assert abs((R[-1] - R[-2]) / (L6 / n6) - 2.0) < 0.04


def _normals(O, om, x):
    return O.measure(om, x)["face_normal"]


@pytest.mark.parametrize("vx", [0.0, 0.37])
def test_uniform_state_surface_speeds(C, O, vx):
    """A uniform B = 0 state moving with V = (vx, 0, 0): S_L = V.n - a at every TS face, S_M = V.n at
    every HP face, S_R = V.n + a at every BS face (a = sqrt(gamma p / rho)), so a node's velocity
    is the mean of velocity_K (V.n +- a) n over its surface faces, and the node moves by its
    radial part (HP behind the Sun: the part along the yz radius)."""
    mesh, _, params, rays = C.shell_fitted(n_edge=4, zone1=False)
    prim = _uniform(mesh, v=(vx, 0.0, 0.0))
    params = dict(params, sglag_TS=0, sglag_HP=0, sglag_BS=0)
    om = O.OracleMesh(mesh)
    F = _fit(O, mesh, rays, kinds=())
    dt = 0.01
    x = mesh.node_xyz
    rc, xn, vel = O.culc_velocity_surface(om, params, F, x, prim, dt)
    assert rc == 0
    n = _normals(O, om, x)
    a = math.sqrt(GAMMA * 0.7 / 1.3)
    kind = F.kind
    vn = vx * n[:, 0]
    ds = np.where(kind == 1, vn - a, np.where(kind == 2, vn, vn + a))
    vk = np.array([0.0, 0.1, 0.1, 0.3])[kind]
    contrib = (vk * ds)[:, None] * n
    for i, v in enumerate(F.sn):
        fs = [f for f in np.nonzero(kind)[0] if v in mesh.face_nodes[f]]
        expect = contrib[fs].mean(axis=0)
        np.testing.assert_allclose(vel[i], expect, rtol=1e-12, atol=1e-14)
        A = x[v]
        B = A.copy()
        if F.sn_kind[i] == 2 and A[0] < 0:
            B[0] = 0.0
        B /= np.linalg.norm(B)
        np.testing.assert_allclose(xn[v], A + (vel[i] @ B) * dt * B, rtol=1e-13, atol=1e-13 * np.linalg.norm(A))
        if F.sn_kind[i] == 2 and A[0] < 0:
            assert xn[v][0] == A[0]
    others = np.setdiff1d(np.arange(mesh.n_nodes), F.sn)
    np.testing.assert_array_equal(xn[others], x[others])


def test_sphere_and_mean_smoothing_are_projections(C, O):
    """With the face terms switched off (move_K = 0 still averages, but only the smoothing term is
    present), the TS and HP velocity times dt / (velocity_K k_K) carries a node onto the sphere
    about its fitted centre through its neighbours' mean distance, along the line from the
    centre; the BS term carries it to the mean of itself and its neighbours."""
    mesh, _, params, rays = C.shell_fitted(n_edge=4, zone1=False)
    x0 = _perturb_surfaces(mesh, rays, seed=7, amp=0.03)
    mesh.node_xyz = x0
    F = _fit(O, mesh, rays)
    params = dict(params, move_TS=0, move_HP=0, move_BS=0, sglag_TS=1, sglag_HP=1, sglag_BS=1)
    om = O.OracleMesh(mesh)
    dt = 0.02
    rc, xn, vel = O.culc_velocity_surface(om, params, F, x0, _uniform(mesh), dt)
    assert rc == 0
    np.testing.assert_array_equal(xn, x0)   # nothing moves with move_K = 0
    coef = {1: 0.1 * 0.01, 2: 0.1 * 0.01, 3: 0.3 * 0.05}
    checked = {1: 0, 2: 0, 3: 0}
    for i, v in enumerate(F.sn):
        k = int(F.sn_kind[i])
        nb = F.nb[i]
        A = x0[v]
        target = A + vel[i] * dt / coef[k]
        if k == 3:
            if len(nb) <= 5:
                assert np.all(vel[i] == 0.0)
                continue
            np.testing.assert_allclose(target, x0[[v] + nb].mean(axis=0), rtol=1e-12)
        else:
            c = F.centre[i]
            r = np.linalg.norm(x0[nb] - c, axis=1).mean()
            np.testing.assert_allclose(np.linalg.norm(target - c), r, rtol=1e-12)
            u = (A - c) / np.linalg.norm(A - c)
            np.testing.assert_allclose((target - c) / np.linalg.norm(target - c), u, atol=1e-12)
        checked[k] += 1
    assert checked[1] > 0 and checked[2] > 0 and checked[3] > 0


def test_sphere_fit_recovers_a_sphere_through_the_origin(T, O):
    """Smooth_head_TS3 solves 2 p.c = |p|^2 without a constant term, which points on a sphere
    through the origin satisfy exactly: a noisy patch of such a sphere gives its centre back to
    the noise level, in the oracle's numpy restatement and in the library's host routine, and
    the two agree."""
    rng = np.random.default_rng(4)
    for trial in range(20):
        c0 = rng.normal(size=3)
        R = np.linalg.norm(c0)
        d = -c0 / R + 0.25 * rng.normal(size=3)
        d /= np.linalg.norm(d)
        pts = []
        for _ in range(9):
            e = d + 0.3 * rng.normal(size=3)
            e /= np.linalg.norm(e)
            pts.append(c0 + R * e)
        pts = np.array(pts) + 1e-6 * R * rng.normal(size=(9, 3))
        co = O.sphere_fit(pts)
        assert np.linalg.norm(co - c0) < 1e-3 * R
        ch = T.sphere_fit(pts)
        np.testing.assert_allclose(ch, co, rtol=1e-9, atol=1e-9 * R)


def test_partial_heliopause_neighbour_count(C, O):
    """macros4: an HP face counts the HP faces of the cells beside its cells[0]; with HP flags on
    one cube panel only, faces on the panel's edge see 3 and corner faces 2, and move no node."""
    ne = 4
    mesh, _, params, rays = C.shell_fitted(n_edge=ne, zone1=False)
    hp = np.nonzero((mesh.face_flags >> 2) & 1)[0]
    ns = 6 * ne * ne + 2
    # keep the HP faces of the first panel (faces whose ray ids all belong to quads 0 .. ne^2-1)
    panel_nodes = set(np.unique(mesh.cell_nodes[:ne * ne] % ns).tolist())
    keep = [f for f in hp if set((mesh.face_nodes[f] % ns).tolist()) <= panel_nodes]
    assert len(keep) == ne * ne
    flags = mesh.face_flags.copy()
    flags[np.setdiff1d(hp, keep)] = 0
    mesh.face_flags = flags
    F = O.OracleFit(mesh, rays)
    moving = set(F.sn_face.tolist())
    kept = sorted(keep)
    interior = [f for f in kept if f in moving]
    assert len(interior) == (ne - 2) ** 2


def test_fitted_step_preserves_free_stream(C, O):
    """The whole moving step (surfaces, rays, re-measure, flux, update) with the exactly
    conservative geometry keeps a uniform state uniform while the surfaces move (zero-gradient
    boundaries)."""
    mesh, _, params, rays = C.shell_fitted(n_edge=4, zone1=False)
    prim = _uniform(mesh, rho=0.9, p=0.4, v=(0.2, -0.1, 0.05))
    # zero-gradient boundaries and no contact flag (w = S_M at the HP would not be the mesh's
    # own face velocity)
    mesh.face_bc = np.where(mesh.face_bc != 0, 4, 0).astype(np.int8)
    mesh.face_flags = (mesh.face_flags & ~np.uint8(2)).astype(np.uint8)
    om = O.OracleMesh(mesh)
    F = _fit(O, mesh, rays)
    params = dict(params, geom_mode=1)
    rc, pf, xf, dth, _ = O.go(om, params, mesh.node_xyz, prim, 6, gcl=True, fit=F)
    assert rc == 0
    assert np.abs(xf - mesh.node_xyz).max() > 1e-6   # the surfaces moved
    np.testing.assert_allclose(pf, prim, rtol=1e-11, atol=1e-12)
