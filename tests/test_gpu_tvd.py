"""Second-order (TVD) face states on the device against the oracle: the neighbour choice of
Setka::Init_TVD (Setka_physics.cpp:648-744), the limited reconstruction of Setka::Snos_on_Gran
(Setka_gran_snos.cpp:186-736) in the face loop, and whole steps of Setka::Go with TVD on.

Faces whose four cells are all outside Zone_1 use only +, -, *, / and square roots, so they are
compared bit for bit; the Zone_1 cases go through spherical frames (acos/atan/sin/cos and pow in
the reference, their algebraic equivalents on the device), compared to 1e-12."""
import numpy as np
import pytest

from helpers import cell_err, cell_errors, var_err

pytestmark = pytest.mark.gpu


def banded_cube(C, n=12, jitter=0.15):
    """Brio-Wu cube whose cells are Zone_2 / Zone_3 / Zone_4 in bands of x, so every zone-transition
    pattern of the reconstruction (2 2 3 3, 2 2 2 3, 2 3 3 3 and the 3/4 ones) occurs."""
    mesh, prim, params = C.cube_brio_wu(n=n, jitter=jitter)
    cc = mesh.node_xyz[mesh.cell_nodes].mean(axis=1)
    mesh.cell_zone = np.where(cc[:, 0] < 0.3, 2, np.where(cc[:, 0] < 0.7, 3, 4)).astype(np.int8)
    rng = np.random.default_rng(11)
    prim = prim.copy()
    prim[:, 1:4] += 0.05 * rng.normal(size=(mesh.n_cells, 3))
    prim[:, 4] *= 1.0 + 0.02 * rng.random(mesh.n_cells)
    return mesh, prim, dict(params, tvd=1)


def small_shell(C, moving=False, n_layers=10):
    mesh, prim, params = C.shell(n_edge=6, n_layers=n_layers, moving=moving, r_zone1_au=80.0, noise=0.01)
    return mesh, prim, dict(params, tvd=1)


@pytest.mark.parametrize("which", ["cube", "shell"])
def test_tvd_neighbours_match_oracle(T, C, O, which):
    mesh, prim, params = banded_cube(C) if which == "cube" else small_shell(C)
    with T.Setka(mesh, params) as s:
        got = s.tvd_cells()
        info = s.tile_info()
    ref = O.OracleMesh(mesh, tvd=True).arrs["face_tvd"]
    np.testing.assert_array_equal(got, ref)
    interior = mesh.face_cells[:, 1] >= 0
    active = interior & (ref[:, 0] >= 0) & (ref[:, 1] >= 0)
    active &= mesh.cell_is_tvd[mesh.face_cells[:, 0]].astype(bool)
    active &= mesh.cell_is_tvd[np.maximum(mesh.face_cells[:, 1], 0)].astype(bool)
    assert info["tvd_faces"] == active.sum() > 0


def test_face_loop_tvd_cube_bit_exact(T, C, O):
    mesh, prim, params = banded_cube(C)
    with T.Setka(mesh, params) as s:
        s.set_state(prim, 0)
        flux, dtmin, br = s.Culc_Gran_Potok(0, 1e-5)
        g = s.geometry(0)
    om = O.OracleMesh(mesh, tvd=True)
    rc, rflux, rdt, rbr = O.face_flux(om, params, prim, g)
    assert rc == 0
    np.testing.assert_array_equal(br, rbr)
    np.testing.assert_array_equal(flux, rflux)
    assert dtmin == min(1e8, rdt.min())
    # second order changed the fluxes
    rc, f1, _, _ = O.face_flux(om, dict(params, tvd=0), prim, g)
    assert np.abs(f1 - rflux).max() > 1e-6


def test_face_loop_tvd_shell(T, C, O):
    mesh, prim, params = small_shell(C)
    with T.Setka(mesh, params) as s:
        s.set_state(prim, 0)
        flux, dtmin, br = s.Culc_Gran_Potok(0, 1e-5)
        g = s.geometry(0)
    rc, rflux, rdt, rbr = O.face_flux(O.OracleMesh(mesh, tvd=True), params, prim, g)
    assert rc == 0
    scale = np.abs(rflux).max(axis=1, keepdims=True) + 1e-300
    assert (np.abs(flux - rflux) / scale).max() < 1e-12
    assert (br != rbr).mean() < 1e-3
    assert abs(dtmin - min(1e8, rdt.min())) <= 1e-12 * dtmin


def _go(T, mesh, prim, params, steps):
    with T.Setka(mesh, params) as s:
        s.set_state(prim, 0)
        st = s.Go(steps)
        return s.get_state(), s.get_nodes(s.current_layer()), st


@pytest.mark.parametrize("precision", [0, 1])
def test_go_tvd_cube(T, C, O, precision):
    """FP64: 50 steps bit for bit.  FP32 Riemann arithmetic: within 1e-5 of the FP64 result."""
    mesh, prim, params = banded_cube(C)
    params = dict(params, precision=precision)
    got, _, st = _go(T, mesh, prim, params, 50)
    rc, ref, _, dth, _ = O.go(O.OracleMesh(mesh, tvd=True), params, mesh.node_xyz, prim, 50, nthreads=8)
    assert rc == 0
    if precision == 0:
        assert st["dt_last"] == dth[-1]
        np.testing.assert_array_equal(got, ref)
    else:
        print("TVD cube, mixed precision, 50 steps: per-cell errors", cell_errors(got, ref))
        assert cell_err(got, ref) < 1e-5


def test_go_tvd_cube_moving_bit_exact(T, C, O):
    """Moving nodes: the reconstruction uses the centres of the step's old layer, which the step
    kernel reads from the previous step's output for cells outside its tile."""
    mesh, prim, params = banded_cube(C, n=14)
    rng = np.random.default_rng(5)
    disp = np.zeros_like(mesh.node_xyz)
    xyz = mesh.node_xyz
    inner = np.all((xyz > 1e-9) & (xyz < 1 - 1e-9), axis=1)
    disp[inner] = 0.01 * rng.normal(size=(inner.sum(), 3))
    mesh.node_disp = disp
    params = dict(params, move_mesh=1, motion_omega=3.0, motion_phase=0.4)
    got, gx, _ = _go(T, mesh, prim, params, 30)
    rc, ref, rx, _, _ = O.go(O.OracleMesh(mesh, tvd=True), params, mesh.node_xyz, prim, 30, nthreads=8, disp=disp)
    assert rc == 0
    np.testing.assert_array_equal(gx, rx)
    np.testing.assert_array_equal(got, ref)


@pytest.mark.parametrize("moving", [False, True])
def test_go_tvd_shell(T, C, O, moving):
    mesh, prim, params = small_shell(C, moving=moving)
    got, xyz, _ = _go(T, mesh, prim, params, 30)
    kw = dict(disp=mesh.node_disp) if moving else {}
    rc, ref, rxyz, _, _ = O.go(O.OracleMesh(mesh, tvd=True), params, mesh.node_xyz, prim, 30, nthreads=8, **kw)
    assert rc == 0
    print("TVD shell, moving %s: per-cell errors %s" % (moving, cell_errors(got, ref)))
    assert cell_err(got, ref) < 1e-12
    # the second-order step differs from the first-order one
    rc, ref1, _, _, _ = O.go(O.OracleMesh(mesh), dict(params, tvd=0), mesh.node_xyz, prim, 30, nthreads=8, **kw)
    assert var_err(ref1, ref) > 1e-8


def test_heliopause_faces_stay_first_order(T, C, O):
    """Faces flagged heliopause (face_flags bit2) reconstruct to second order only with
    Snos_on_HP (Setka_gran_snos.cpp:188-189)."""
    mesh, prim, params = banded_cube(C, n=10)
    fx = mesh.node_xyz[mesh.face_nodes].mean(axis=1)[:, 0]
    flags = np.zeros(mesh.n_faces, np.uint8)
    flags[np.abs(fx - 0.5) < 0.08] = 4
    mesh.face_flags = flags
    assert flags.any()
    for hp in (0, 1):
        p = dict(params, snos_on_hp=hp)
        with T.Setka(mesh, p) as s:
            s.set_state(prim, 0)
            flux, _, _ = s.Culc_Gran_Potok(0, 1e-5)
            g = s.geometry(0)
        rc, rflux, _, _ = O.face_flux(O.OracleMesh(mesh, tvd=True), p, prim, g)
        assert rc == 0
        np.testing.assert_array_equal(flux, rflux)
    rc, f0, _, _ = O.face_flux(O.OracleMesh(mesh, tvd=True), dict(params, snos_on_hp=0), prim, g)
    rc, f1, _, _ = O.face_flux(O.OracleMesh(mesh, tvd=True), dict(params, snos_on_hp=1), prim, g)
    hpf = flags == 4
    assert np.abs(f0[hpf] - f1[hpf]).max() > 0.0
    np.testing.assert_array_equal(f0[~hpf], f1[~hpf])


def test_go_tvd_shell_zone_bands(T, C, O):
    """Radial bands Zone_1 | Zone_2 | Zone_3 | Zone_4: every zone pattern of the reconstruction table
    (the three Zone_1/Zone_2 transitions with spherical variables, the six plain transitions) occurs
    on the faces; 20 steps against the oracle at 1e-12 (Zone_1 trig)."""
    mesh, prim, params = small_shell(C, n_layers=12)
    r_in = np.linalg.norm(mesh.node_xyz, axis=1)[mesh.cell_nodes].min(axis=1)
    layer = np.searchsorted(np.unique(np.round(r_in, 9)), np.round(r_in, 9))
    mesh.cell_zone = (1 + np.minimum(layer // 3, 3)).astype(np.int8)
    om = O.OracleMesh(mesh, tvd=True)
    tvd = om.arrs["face_tvd"]
    ok = (mesh.face_cells[:, 1] >= 0) & (tvd[:, 0] >= 0) & (tvd[:, 1] >= 0)
    z = mesh.cell_zone
    pats = {(int(z[tvd[f, 0]]), int(z[mesh.face_cells[f, 0]]), int(z[mesh.face_cells[f, 1]]), int(z[tvd[f, 1]]))
            for f in np.nonzero(ok)[0]}
    for p in ((1, 1, 1, 2), (1, 1, 2, 2), (1, 2, 2, 2), (2, 2, 2, 3), (2, 2, 3, 3), (2, 3, 3, 3),
              (3, 3, 3, 4), (3, 3, 4, 4), (3, 4, 4, 4), (1, 1, 1, 1)):
        assert p in pats, p
    got, _, _ = _go(T, mesh, prim, params, 20)
    rc, ref, _, _, _ = O.go(om, params, mesh.node_xyz, prim, 20, nthreads=8)
    assert rc == 0
    print("TVD zone-banded shell: per-cell errors", cell_errors(got, ref))
    assert cell_err(got, ref) < 1e-12
