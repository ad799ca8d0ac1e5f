"""titan_mk_go (Setka::MK_go / MK_fly_immit, Setka_MK.cpp:1225-2316) on the device against the
oracle's or_mk_go and against what the kinetic theory fixes.  The particle counts and weights are
integer / exact work and match the oracle exactly; the flights use the same random stream per
particle, so most particles follow the oracle's path, but the device's transcendental functions
differ from the C library's in the last bits and a rejection or exit test can then go the other
way: the densities are compared statistically (MC noise ~ 1/sqrt(particles per cell))."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ME = 1.0 / 1836.15
A2 = 0.102046


def _cube(C, n=6):
    mesh, prim, params = C.cube_brio_wu(n=n, jitter=0.1)
    mesh.face_bc = np.where(mesh.face_cells[:, 1] < 0, C.BC_OUTER_SOFT, C.BC_INTERIOR).astype(np.int8)
    return mesh, prim, params


def _lism(C, mesh):
    prim = np.zeros((mesh.n_cells, 8))
    prim[:, 0] = 1.0
    prim[:, 1] = C.V_INF
    prim[:, 4] = 1.0 / (1.0 + ME)
    return prim


def _run(T, mesh, prim, params, n_per_face, seed, kn):
    with T.Setka(mesh, dict(params, par_a_2=A2, par_Kn=kn)) as s:
        s.set_state(prim, 0)
        return s.MK_go(n_per_face, seed=seed, layer=0)


def test_free_flight_matches_the_oracle_and_the_closed_form(T, C, O):
    mesh, prim, params = _cube(C)
    d, st = _run(T, mesh, prim, params, 200, 7, 1e30)
    od, ost = O.mk_go(O.OracleMesh(mesh), mesh.node_xyz, prim, C.V_INF, A2, 1e30, 200, seed=7)
    assert st["particles"] == ost["particles"] and st["exchanges"] == 0
    assert st["left"] + st["births_dropped"] == st["particles"]
    assert abs(st["births_dropped"] - ost["births_dropped"]) <= max(3, ost["births_dropped"] // 100)
    assert abs(st["segments"] - ost["segments"]) < 0.01 * ost["segments"]
    assert abs(d[:, 0].mean() - 1.0) < 0.012 and np.abs(d[:, 0] - 1.0).max() < 0.15
    # same random stream per particle: the cell densities agree far inside the MC noise
    assert np.abs(d[:, 0] - od[:, 0]).max() < 0.03
    assert abs(d[:, 0].mean() - od[:, 0].mean()) < 2e-3
    np.testing.assert_allclose(d[:, 4], d[:, 0], rtol=1e-12)


@pytest.mark.parametrize("kn", [1.0, 0.3])
def test_detailed_balance_and_oracle(T, C, O, kn):
    mesh, _, params = _cube(C)
    prim = _lism(C, mesh)
    d, st = _run(T, mesh, prim, params, 200, 11, kn)
    od, ost = O.mk_go(O.OracleMesh(mesh), mesh.node_xyz, prim, C.V_INF, A2, kn, 200, seed=11)
    assert st["particles"] == ost["particles"] and st["lost"] == 0
    assert abs(st["exchanges"] - ost["exchanges"]) < 0.02 * ost["exchanges"]
    assert abs(d[:, 0].mean() - 1.0) < 0.012 and np.abs(d[:, 0] - 1.0).max() < 0.15
    np.testing.assert_allclose(d[:, 1:].sum(axis=1), d[:, 0], rtol=1e-12)
    assert np.all(d[:, 1] == 0.0) and np.all(d[:, 2] == 0.0)
    # sort shares: the device's exchanged fraction agrees with the oracle's within MC noise
    assert abs(d[:, 3].sum() / d[:, 0].sum() - od[:, 3].sum() / od[:, 0].sum()) < 0.01
    assert np.abs(d[:, 0] - od[:, 0]).max() < 0.12


def test_repeat_is_reproducible_to_rounding(T, C):
    mesh, _, params = _cube(C, n=4)
    prim = _lism(C, mesh)
    d1, s1 = _run(T, mesh, prim, params, 100, 5, 1.0)
    d2, s2 = _run(T, mesh, prim, params, 100, 5, 1.0)
    assert s1 == s2
    np.testing.assert_allclose(d1, d2, rtol=1e-12, atol=1e-15)
    d3, _ = _run(T, mesh, prim, params, 100, 6, 1.0)
    assert np.abs(d3 - d1).max() > 0.0            # another seed, another sample


def test_fitted_shell_against_the_oracle(T, C, O):
    """The fitted-shell test mesh with its solar-wind state: the inflow enters through the outer
    sphere (bc 2, 3) and crosses the heliosheath and the wind.  With the same random stream per
    particle the device's volume-weighted densities lie closer to the oracle's than two oracle
    samples of different seeds lie to each other (the MC noise of the estimate)."""
    mesh, prim, params, rays = C.shell_fitted(zone1=False)
    d, st = _run(T, mesh, prim, params, 100, 13, 40.0906)
    od, ost = O.mk_go(O.OracleMesh(mesh), mesh.node_xyz, prim, C.V_INF, A2, 40.0906, 100, seed=13)
    assert st["particles"] == ost["particles"] > 0
    assert st["lost"] == 0 and ost["lost"] == 0
    assert st["left"] + st["births_dropped"] == st["particles"]
    assert abs(st["sun_passes"] - ost["sun_passes"]) < 0.1 * ost["sun_passes"] + 10
    assert abs(st["exchanges"] - ost["exchanges"]) < 0.03 * ost["exchanges"]
    o2, _ = O.mk_go(O.OracleMesh(mesh), mesh.node_xyz, prim, C.V_INF, A2, 40.0906, 100, seed=14)
    vol = O.measure(O.OracleMesh(mesh), mesh.node_xyz)["cell_volume"]
    for k in (0, 3, 4):
        same = (np.abs(d[:, k] - od[:, k]) * vol).sum()
        noise = (np.abs(o2[:, k] - od[:, k]) * vol).sum()
        assert same < 0.5 * noise, (k, same, noise)


def test_pass_through_the_sun(T, C, O):
    """The Sun shell (LISM free flight around the inner sphere): the device carries the particles
    across the sphere as the oracle does -- no particle lost, the pass counts alike, the density 1
    past the first layer and in the cone behind the Sun, the cell densities close to the oracle's
    (same random streams)."""
    mesh, prim, params = C.sun_shell()
    d, st = _run(T, mesh, prim, params, 1000, 5, 1e30)
    od, ost = O.mk_go(O.OracleMesh(mesh), mesh.node_xyz, prim, C.V_INF, A2, 1e30, 1000, seed=5, nthreads=8)
    assert st["particles"] == ost["particles"] and st["lost"] == 0 and ost["lost"] == 0
    assert st["left"] + st["births_dropped"] == st["particles"]
    assert abs(st["sun_passes"] - ost["sun_passes"]) < 0.01 * ost["sun_passes"] + 10
    from test_oracle_mk import _sun_layers
    lay = _sun_layers(O, mesh, d)
    assert np.abs(lay[:, 0] - 1.0).max() < 0.012, lay
    assert np.abs(lay[:, 1] - 1.0).max() < 0.03, lay
    assert np.abs(d[:, 0] - od[:, 0]).max() < 0.03


def test_bad_arguments(T, C):
    mesh, prim, params = _cube(C, n=4)
    with T.Setka(mesh, params) as s:
        s.set_state(prim, 0)
        for kw in (dict(n_per_face=0), dict(n_per_face=10, max_steps=0), dict(n_per_face=10, layer=2)):
            with pytest.raises(T.TitanError) as e:
                s.MK_go(**kw)
            assert e.value.code == T.TITAN_E_ARG


def test_no_outer_boundary_gives_nothing(T, C):
    """A mesh without outer-boundary faces injects no particle: zero densities and counts."""
    mesh, prim, params = C.cube_brio_wu(n=4, jitter=0.1)   # every boundary face zero-gradient (bc 4)
    with T.Setka(mesh, params) as s:
        s.set_state(prim, 0)
        d, st = s.MK_go(50)
    assert np.all(d == 0.0) and all(v == 0 for v in st.values())


def test_second_call_same_context(T, C):
    """The injection list is found once per context: a second call with the same seed repeats the
    first (counts exactly, densities to rounding)."""
    mesh, _, params = _cube(C, n=4)
    prim = _lism(C, mesh)
    with T.Setka(mesh, dict(params, par_Kn=1.0)) as s:
        s.set_state(prim, 0)
        d1, s1 = s.MK_go(60, seed=4, layer=0)
        d2, s2 = s.MK_go(60, seed=4, layer=0)
    assert s1 == s2
    np.testing.assert_allclose(d1, d2, rtol=1e-12, atol=1e-15)


def test_moved_layer(T, C, O):
    """After moving steps, MK_go of the current layer flies through that layer's nodes: the device
    agrees with the oracle run on the moved nodes and the stepped state (same seed closer than
    another seed), and the particle counts match exactly."""
    mesh, _, params = _cube(C, n=5)
    prim = _lism(C, mesh)
    rng = np.random.default_rng(8)
    xyz = mesh.node_xyz
    inside = np.all((xyz > 1e-9) & (xyz < 1 - 1e-9), axis=1)
    disp = np.zeros_like(xyz)
    disp[inside] = 0.02 * rng.normal(size=(inside.sum(), 3))
    mesh.node_disp = disp
    p = dict(params, move_mesh=1, motion_omega=3.0, motion_phase=0.4, par_Kn=1.0)
    with T.Setka(mesh, p) as s:
        s.set_state(prim, 0)
        s.Go(3)
        lay = s.current_layer()
        x_now = s.get_nodes(lay)
        p_now = s.get_state(lay)
        d, st = s.MK_go(100, seed=21)
    assert np.abs(x_now - mesh.node_xyz).max() > 1e-6
    om = O.OracleMesh(mesh)
    od, ost = O.mk_go(om, x_now, p_now, C.V_INF, A2, 1.0, 100, seed=21)
    o2, _ = O.mk_go(om, x_now, p_now, C.V_INF, A2, 1.0, 100, seed=22)
    assert st["particles"] == ost["particles"]
    vol = O.measure(om, x_now)["cell_volume"]
    same = (np.abs(d[:, 0] - od[:, 0]) * vol).sum()
    noise = (np.abs(o2[:, 0] - od[:, 0]) * vol).sum()
    assert same < 0.5 * noise, (same, noise)


def test_helium_split_matches_the_oracle(T, C, O):
    """With the passive scalars carried, the flight's protons are the thermal part of
    Plasma_components_1 at the cell's type (Setka_MK.cpp:1944-1966): rho_He = 0.3 rho and p set so
    that cp stays 1, on both sides; the exchange count follows the oracle's, and falls to 0.7 of
    the helium-free count as rho_Th does."""
    mesh, _, params = _cube(C, n=5)
    prim = _lism(C, mesh)
    he = 0.3
    al = np.where(np.asarray(mesh.cell_zone) <= 2, 2, 1)
    r_th = -(ME * al * he + 4.0 * (-1.0 + he)) / (4.0 * (1.0 + ME))
    ph = prim.copy()
    ph[:, 4] = (r_th / 2.0) * (8.0 + (-7.0 + al + (1.0 - al) * ME) * he) / (4.0 - (4.0 + al * ME) * he)
    scal = np.zeros((mesh.n_cells, 2))
    scal[:, 1] = he
    with T.Setka(mesh, dict(params, par_a_2=A2, par_Kn=0.3, passive_scalars=1)) as s:
        s.set_state(ph, 0)
        s.set_scalars(scal, 0)
        d, st = s.MK_go(150, seed=22, layer=0)
        s.set_scalars(np.zeros_like(scal), 0)
        s.set_state(prim, 0)
        d0, st0 = s.MK_go(150, seed=21, layer=0)
    od, ost = O.mk_go(O.OracleMesh(mesh), mesh.node_xyz, ph, C.V_INF, A2, 0.3, 150, seed=22, nthreads=4, scal=scal)
    assert st["particles"] == ost["particles"] and st["lost"] == 0
    assert abs(st["exchanges"] - ost["exchanges"]) < 0.01 * ost["exchanges"]
    assert abs(d[:, 0].mean() - od[:, 0].mean()) < 2e-3
    assert abs(st["exchanges"] / st0["exchanges"] - r_th.mean() * (1.0 + ME)) < 0.02
    # Q = 0 carried: determ_zone sends the exchanged atoms of these type-3 cells to sort 2, not 3
    assert d[:, 3].sum() == 0 and od[:, 3].sum() == 0 and d[:, 2].sum() > 0
    assert abs(d[:, 2].sum() / d[:, 0].sum() - od[:, 2].sum() / od[:, 0].sum()) < 0.01
