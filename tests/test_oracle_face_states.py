"""Pins of the oracle's face states (Setka::Snos_on_Gran, first order and boundary faces,
Setka_gran_snos.cpp:3-180,737-1071) and of the prescribed node motion.

The Zone_1 rescaling is pinned by what it is built to do: a spherically symmetric wind
(rho ~ r^-2, p ~ r^-2gamma, radial V, radial B ~ r^-2, azimuthal B ~ r^-1) sampled at cell
centres must arrive at every face as the same wind evaluated at the face centre, in both
spherical frames the reference switches between."""
import math

import numpy as np


def _zone1_shell(C, O):
    mesh, _, params = C.shell(n_edge=3, n_layers=4, r_zone1_au=1e9)
    om = O.OracleMesh(mesh)
    g = O.measure(om, mesh.node_xyz)
    return mesh, params, om, g


def _wind(x, gamma, v_tan=0.0, b_tan=0.0, axis=2):
    """The symmetric wind at points x; tangential parts are azimuthal about `axis`."""
    r = np.linalg.norm(x, axis=-1)
    rh = x / r[..., None]
    e_ax = np.zeros(3)
    e_ax[axis] = 1.0
    phi = np.cross(e_ax, x)
    nphi = np.linalg.norm(phi, axis=-1)[..., None]
    phi = np.divide(phi, nphi, out=np.zeros_like(phi), where=nphi > 0)
    out = np.zeros(x.shape[:-1] + (8,))
    out[..., 0] = 3.0 / r ** 2
    out[..., 4] = 2.0 * r ** (-2.0 * gamma)
    out[..., 1:4] = 5.0 * rh + v_tan * phi
    out[..., 5:8] = 0.7 / r[..., None] ** 2 * rh + (b_tan / r)[..., None] * phi
    return out


def _check_faces(mesh, params, om, g, prim, faces, expect_fn):
    scale = np.abs(prim).max(axis=0) + 1e-300
    for f in faces:
        rc, qL, qR = O_face(om, params, prim, g, f)
        exp = expect_fn(g["face_center"][f])
        assert rc == 0
        for q in (qL, qR):
            assert np.all(np.abs(q - exp) <= 1e-13 * scale), (f, q, exp)


O_face = None


def test_zone1_radial_wind_is_carried_to_the_face(C, O):
    global O_face
    O_face = O.face_states
    mesh, params, om, g = _zone1_shell(C, O)
    gamma = params["gamma"]
    prim = _wind(g["cell_center"], gamma)
    interior = np.nonzero(mesh.face_cells[:, 1] >= 0)[0]
    _check_faces(mesh, params, om, g, prim, interior, lambda G: _wind(G, gamma))


def test_zone1_tangential_parts_in_both_frames(C, O):
    """Azimuthal V is carried unscaled and azimuthal B scaled by r_cell/r_face; the frame is
    polar about z when the face's polar angle lies in (pi/9, 8pi/9), else about x."""
    global O_face
    O_face = O.face_states
    mesh, params, om, g = _zone1_shell(C, O)
    gamma = params["gamma"]
    G = g["face_center"]
    theta = np.arccos(G[:, 2] / np.linalg.norm(G, axis=1))
    interior = mesh.face_cells[:, 1] >= 0
    equatorial = (theta > math.pi / 9) & (theta < 8 * math.pi / 9)
    for axis, sel in ((2, equatorial), (0, ~equatorial)):
        faces = np.nonzero(interior & sel)[0]
        assert faces.size > 0
        prim = _wind(g["cell_center"], gamma, v_tan=0.4, b_tan=0.9, axis=axis)
        # only faces whose two cells are away from the frame axis (no azimuth ambiguity)
        cc = g["cell_center"][mesh.face_cells[faces]]
        off_axis = np.all(np.linalg.norm(np.delete(cc, axis, axis=2), axis=2) > 1e-6, axis=1)
        _check_faces(mesh, params, om, g, prim, faces[off_axis],
                     lambda x: _wind(x, gamma, v_tan=0.4, b_tan=0.9, axis=axis))


def test_boundary_rules(C, O):
    """Outer-soft suction fix (Setka_gran_snos.cpp:18-50), hard faces (fixed state), zero
    gradient (copy), on the faces of an axis-aligned cube."""
    mesh, _, params = C.cube_brio_wu(n=3, jitter=0.0)
    om0 = O.OracleMesh(mesh)
    g = O.measure(om0, mesh.node_xyz)
    n = g["face_normal"]
    bnd = mesh.face_cells[:, 1] < 0
    f_lo = int(np.nonzero(bnd & (n[:, 0] < -0.99))[0][0])   # x = 0 side, n = -x
    f_hi = int(np.nonzero(bnd & (n[:, 0] > 0.99))[0][0])    # x = 1 side, n = +x
    f_y = int(np.nonzero(bnd & (np.abs(n[:, 1]) > 0.99))[0][0])
    v_inf = params["v_inf"]
    bc = mesh.face_bc.copy()
    bstate = np.zeros((mesh.n_faces, 8))
    bstate[f_y] = np.arange(1, 9)
    bc[f_lo] = bc[f_hi] = 3
    bc[f_y] = 2
    mesh.face_bc = bc
    mesh.face_bstate = bstate
    om = O.OracleMesh(mesh)
    base = np.array([1.0, 0.0, 0.2, 0.3, 1.5, 0.1, 0.2, 0.3])

    def states(f, v):
        prim = np.tile(base, (mesh.n_cells, 1))
        prim[:, 1:4] = v
        rc, qL, qR = O.face_states(om, params, prim, g, f)
        assert rc == 0
        return prim[mesh.face_cells[f, 0]], qL, qR

    # n_x < -0.95 and Vx > V_inf/7: Vx -> V_inf/5, other components kept, right = left
    s, qL, qR = states(f_lo, [0.0, 0.2, 0.3])
    assert qL[1] == v_inf / 5.0 and np.array_equal(qL[2:], s[2:]) and np.array_equal(qL, qR)
    # n_x < -0.95 and Vx below V_inf/7: untouched
    s, qL, qR = states(f_lo, [v_inf, 0.2, 0.3])
    assert np.array_equal(qL, s) and np.array_equal(qR, s)
    # elsewhere, inflow (V.n < 0) is replaced by 0.1 n
    s, qL, qR = states(f_hi, [-0.3, 0.2, 0.3])
    assert np.array_equal(qL[1:4], 0.1 * n[f_hi]) and np.array_equal(qL, qR)
    # outflow untouched
    s, qL, qR = states(f_hi, [0.3, 0.2, 0.3])
    assert np.array_equal(qL, s) and np.array_equal(qR, s)
    # hard face: right state is the face's fixed state
    s, qL, qR = states(f_y, [0.3, 0.2, 0.3])
    assert np.array_equal(qL, s) and np.array_equal(qR, bstate[f_y])
    # zero gradient: copy
    f_z = int(np.nonzero(bnd & (np.abs(n[:, 2]) > 0.99))[0][0])
    s, qL, qR = states(f_z, [0.3, 0.2, 0.3])
    assert np.array_equal(qL, s) and np.array_equal(qR, s)


def test_motion_rigid_translation(O):
    """A constant displacement field moves every node by dt*omega*cos(omega*t + phase)*D."""
    rng = np.random.default_rng(4)
    x = rng.normal(size=(50, 3))
    D = np.tile([0.3, -1.2, 0.5], (50, 1))
    dt, t, om, ph = 1e-3, 0.37, 100.0, 1.1
    y = O.motion(x, D, dt, t, om, ph)
    d = y - x
    assert np.allclose(d, d[0], rtol=0, atol=1e-15)
    assert np.allclose(d[0], dt * om * math.cos(om * t + ph) * D[0], rtol=1e-12, atol=0)
