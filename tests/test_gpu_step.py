"""Parity of the device geometry, face loop and full step with the oracle (Setka::Go)."""
import hashlib

import numpy as np
import pytest

from helpers import cell_err, cell_errors, node_err, rel_err

pytestmark = pytest.mark.gpu


def small_shell(C, moving=False, n_edge=6, n_layers=10, **kw):
    kw.setdefault("noise", 0.01)
    return C.shell(n_edge=n_edge, n_layers=n_layers, moving=moving, r_zone1_au=80.0, **kw)


@pytest.mark.parametrize("gcl", [False, True])
def test_geometry_bit_exact(T, C, O, gcl):
    for mesh, prim, params in (C.cube_brio_wu(n=12), small_shell(C)):
        params = dict(params, geom_mode=T.GEOM_GCL if gcl else T.GEOM_REF)
        with T.Setka(mesh, params) as s:
            g = s.geometry(0)
        ref = O.measure(O.OracleMesh(mesh), mesh.node_xyz, gcl=gcl)
        for k in ref:
            np.testing.assert_array_equal(g[k], ref[k], err_msg=k)


def test_face_loop_cube_bit_exact(T, C, O):
    mesh, prim, params = C.cube_brio_wu()
    with T.Setka(mesh, params) as s:
        s.set_state(prim, 0)
        flux, dtmin, br = s.Culc_Gran_Potok(0, 1e-5)
        g = s.geometry(0)
    rc, rflux, rdt, rbr = O.face_flux(O.OracleMesh(mesh), params, prim, g)
    assert rc == 0
    np.testing.assert_array_equal(br, rbr)
    np.testing.assert_array_equal(flux, rflux)
    assert dtmin == min(1e8, rdt.min())


def test_face_loop_shell_zone1(T, C, O):
    """Faces touching Zone_1 cells use acos/atan/sin/cos/pow: GPU libm vs glibc differ in the
    last ulp there, so those faces are compared to 1e-12 relative; all others bit for bit."""
    mesh, prim, params = small_shell(C)
    with T.Setka(mesh, params) as s:
        s.set_state(prim, 0)
        flux, dtmin, br = s.Culc_Gran_Potok(0, 1e-5)
        g = s.geometry(0)
    rc, rflux, rdt, rbr = O.face_flux(O.OracleMesh(mesh), params, prim, g)
    assert rc == 0
    np.testing.assert_array_equal(br, rbr)
    z1 = mesh.cell_zone[mesh.face_cells[:, 0]] == 1
    fb = mesh.face_cells[:, 1]
    z1 |= (fb >= 0) & (mesh.cell_zone[np.maximum(fb, 0)] == 1)
    assert z1.any() and (~z1).any()
    np.testing.assert_array_equal(flux[~z1], rflux[~z1])
    scale = np.abs(rflux[z1]).max(axis=1, keepdims=True)
    assert (np.abs(flux[z1] - rflux[z1]) / scale).max() < 1e-12
    assert abs(dtmin - min(1e8, rdt.min())) <= 1e-12 * dtmin


def test_go_cube_brio_wu_100_steps_bit_exact(T, C, O):
    mesh, prim, params = C.cube_brio_wu()
    with T.Setka(mesh, params) as s:
        s.set_state(prim, 0)
        st = s.Go(100)
        got = s.get_state()
    rc, ref, _, dth, _ = O.go(O.OracleMesh(mesh), params, mesh.node_xyz, prim, 100, nthreads=8)
    assert rc == 0
    assert st["dt_last"] == dth[-1]
    np.testing.assert_array_equal(got, ref)


def _run_gpu(T, mesh, prim, params, steps):
    with T.Setka(mesh, params) as s:
        s.set_state(prim, 0)
        s.Go(steps)
        return s.get_state(), s.get_nodes(s.current_layer())


def test_go_shell_fixed(T, C, O):
    mesh, prim, params = small_shell(C)
    got, _ = _run_gpu(T, mesh, prim, params, 30)
    rc, ref, _, _, _ = O.go(O.OracleMesh(mesh), params, mesh.node_xyz, prim, 30, nthreads=8)
    assert rc == 0
    # Zone_1 faces use acos/atan/sin/cos/pow (GPU libm vs glibc: last-ulp differences); per cell,
    # each conserved variable relative to the cell's own value
    print("shell fixed, 30 steps: per-cell errors", cell_errors(got, ref))
    assert cell_err(got, ref) < 1e-12


def test_go_shell_moving(T, C, O):
    mesh, prim, params = small_shell(C, moving=True)
    got, xyz = _run_gpu(T, mesh, prim, params, 30)
    rc, ref, rxyz, _, _ = O.go(O.OracleMesh(mesh), params, mesh.node_xyz, prim, 30, nthreads=8,
                               disp=mesh.node_disp)
    assert rc == 0
    assert np.abs(xyz - mesh.node_xyz).max() > 0.0
    # the lagged dt inherits last-ulp differences of the Zone_1 faces, and the nodes move by dt
    print("shell moving, 30 steps: per-cell errors", cell_errors(got, ref), "nodes", node_err(xyz, rxyz))
    assert node_err(xyz, rxyz) < 1e-14
    assert cell_err(got, ref) < 1e-12


def test_go_cube_moving_bit_exact(T, C, O):
    """No transcendental functions on this path: the moving-mesh step (motion law, REF geometry,
    face velocities) must agree bit for bit."""
    mesh, prim, params = C.cube_brio_wu(n=16)
    rng = np.random.default_rng(3)
    disp = np.zeros_like(mesh.node_xyz)
    xyz = mesh.node_xyz
    inner = np.all((xyz > 1e-9) & (xyz < 1 - 1e-9), axis=1)
    disp[inner] = 0.01 * rng.normal(size=(inner.sum(), 3))
    mesh.node_disp = disp
    params = dict(params, move_mesh=1, motion_omega=3.0, motion_phase=0.4)
    got, gx = _run_gpu(T, mesh, prim, params, 40)
    rc, ref, rx, _, _ = O.go(O.OracleMesh(mesh), params, mesh.node_xyz, prim, 40, nthreads=8, disp=disp)
    assert rc == 0
    np.testing.assert_array_equal(gx, rx)
    np.testing.assert_array_equal(got, ref)


def test_repeatable_bit_for_bit(T, C):
    mesh, prim, params = small_shell(C, moving=True)
    h = []
    for _ in range(2):
        got, _ = _run_gpu(T, mesh, prim, params, 20)
        h.append(hashlib.sha256(got.tobytes()).hexdigest())
    assert h[0] == h[1]


def test_gcl_free_stream(T, C):
    """Exactly conservative geometry: a uniform state on the moving shell stays uniform."""
    mesh, prim, params = small_shell(C, moving=True)
    u = np.array([1.0, 0.3, -0.2, 0.1, 2.0, 0.5, 0.4, -0.3])
    prim = np.tile(u, (mesh.n_cells, 1))
    mesh.face_bc[mesh.face_bc != T.BC_INTERIOR] = T.BC_ZERO_GRADIENT
    mesh.face_method[:] = 3
    mesh.cell_zone[:] = 3
    params = dict(params, geom_mode=T.GEOM_GCL)
    got, xyz = _run_gpu(T, mesh, prim, params, 20)
    assert np.abs(xyz - mesh.node_xyz).max() > 1e-4
    assert rel_err(got, np.broadcast_to(u, got.shape), 1e-30).max() < 1e-11


@pytest.mark.parametrize("precision", [0, 1])
def test_nan_state_stops_where_the_reference_stops(T, C, O, precision):
    """A NaN density reaches chlld's fatal branch on the faces of its cell before any new density is
    formed; the reference's face loop runs before its cell loop (Setka_physics.cpp:1634-1671) and
    exits at the first failing face, so the step stops with TITAN_E_RIEMANN and the smallest failing
    face id -- whichever workgroup runs first (errors are ranked, face loop before cell loop)."""
    mesh, prim, params = C.cube_brio_wu(n=8)
    prim = prim.copy()
    bad_cell = 37
    prim[bad_cell, 0] = np.nan
    om = O.OracleMesh(mesh)
    g = O.measure(om, mesh.node_xyz)
    failing = []
    for f in np.nonzero((mesh.face_cells == bad_cell).any(axis=1))[0]:
        a, b = mesh.face_cells[f]
        qL = prim[a]
        qR = prim[b] if b >= 0 else prim[a]
        *_, rc = O.chlld_batch(int(mesh.face_method[f]), g["face_normal"][f][None, :], qL[None, :], qR[None, :])
        if rc != 0:
            failing.append(int(f))
    assert failing, "a NaN density reaches chlld's fatal branch"
    for _ in range(2):
        with T.Setka(mesh, dict(params, precision=precision)) as s:
            s.set_state(prim, 0)
            with pytest.raises(T.TitanError) as e:
                s.Go(1)
        assert e.value.code == T.TITAN_E_RIEMANN
        assert e.value.index == min(failing)


@pytest.mark.parametrize("precision", [0, 1])
def test_tile_layout_does_not_change_the_result(T, C, O, precision, monkeypatch):
    """Faces on a tile border are solved by both tiles from the same inputs with the same
    arithmetic, so the step's bits do not depend on where the tiles are cut (tile size set through
    TITAN_TILE_CELLS at titan_create): 20 moving-shell steps, FP64 and FP32-mixed."""
    mesh, prim, params = small_shell(C, moving=True)
    params = dict(params, precision=precision)
    runs = []
    for cells in ("128", "48", "16"):
        monkeypatch.setenv("TITAN_TILE_CELLS", cells)
        with T.Setka(mesh, params) as s:
            s.set_state(prim, 0)
            s.Go(20)
            runs.append((s.get_state(), s.tile_info()["tiles"]))
    assert runs[0][1] < runs[1][1] < runs[2][1]
    for got, _ in runs[1:]:
        np.testing.assert_array_equal(got, runs[0][0])
