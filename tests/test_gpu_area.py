"""titan_go_area -- Setka::Go(is_inner_area, steps) with the reference's per-call dt restart -- against
the oracle's Go on the same area (params area 1 / 2), bit for bit on the cube (fixed and moving)
and on the fitted-surface shell without Zone_1; chained calls restart dt; a NaN deep in the
frozen region does not stop an area step."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _cube(C, moving=False):
    mesh, prim, params = C.cube_brio_wu(n=10, jitter=0.1)
    cc = mesh.node_xyz[mesh.cell_nodes].mean(axis=1)
    inner = (cc[:, 0] < 0.4).astype(np.uint8)
    if moving:
        rng = np.random.default_rng(3)
        xyz = mesh.node_xyz
        inside = np.all((xyz > 1e-9) & (xyz < 1 - 1e-9), axis=1)
        disp = np.zeros_like(xyz)
        disp[inside] = 0.01 * rng.normal(size=(inside.sum(), 3))
        mesh.node_disp = disp
        params = dict(params, move_mesh=1, motion_omega=3.0, motion_phase=0.4)
    return mesh, prim, params, inner


@pytest.mark.parametrize("moving", [False, True])
@pytest.mark.parametrize("is_inner", [True, False])
def test_go_area_cube_bit_exact(T, C, O, moving, is_inner):
    mesh, prim, params, inner = _cube(C, moving)
    with T.Setka(mesh, params) as s:
        s.set_inner_area(inner)
        s.set_state(prim, 0)
        s.Go(7, is_inner_area=is_inner)
        got = s.get_state()
        gx = s.get_nodes(s.current_layer())
    om = O.OracleMesh(mesh)
    om.set_inner(inner)
    rc, pf, xf, _, _ = O.go(om, dict(params, area=1 if is_inner else 2), mesh.node_xyz, prim, 7,
                            disp=mesh.node_disp)
    assert rc == 0
    np.testing.assert_array_equal(got, pf)
    np.testing.assert_array_equal(gx, xf)


def test_chained_calls_restart_dt(T, C, O):
    """Go(outer, 4) then Go(inner, 2), as TITAN.cpp:121-122 alternates them: each call starts
    again from dt = 1e-5, so the device state equals two separate oracle runs chained."""
    mesh, prim, params, inner = _cube(C)
    with T.Setka(mesh, params) as s:
        s.set_inner_area(inner)
        s.set_state(prim, 0)
        s.Go(4, is_inner_area=False)
        st = s.Go(2, is_inner_area=True)
        got = s.get_state()
    om = O.OracleMesh(mesh)
    om.set_inner(inner)
    rc, p1, x1, _, _ = O.go(om, dict(params, area=2), mesh.node_xyz, prim, 4)
    rc2, p2, _, d2, _ = O.go(om, dict(params, area=1), x1, p1, 2)
    assert rc == rc2 == 0
    assert d2[0] == 1e-5 and st["dt_last"] == d2[1]
    np.testing.assert_array_equal(got, p2)


def test_go_area_fitted_shell(T, C, O):
    mesh, prim, params, rays = C.shell_fitted(zone1=False, noise=0.01)
    inner = C.inner_cells_from_rays(mesh, rays)
    assert 0 < inner.sum() < mesh.n_cells
    F = O.OracleFit(mesh, rays)
    for k in (1, 2, 3):
        F.smooth(k, mesh.node_xyz)
    for i in range(len(F.sn)):
        if F.sn_kind[i] != 3 and F.nb[i]:
            F.centre[i] = T.sphere_fit(mesh.node_xyz[F.cand[i]])
    with T.Setka(mesh, params) as s:
        s.set_rays(rays)
        s.Smooth_head_TS3()
        s.Smooth_head_HP3()
        s.Find_Yzel_Sosed_for_BS()
        s.set_inner_area(inner)
        s.set_state(prim, 0)
        s.Go(4, is_inner_area=False)
        got = s.get_state()
        gx = s.get_nodes(s.current_layer())
    om = O.OracleMesh(mesh)
    om.set_inner(inner)
    rc, pf, xf, _, _ = O.go(om, dict(params, area=2), mesh.node_xyz, prim, 4, fit=F)
    assert rc == 0
    np.testing.assert_array_equal(gx, xf)
    np.testing.assert_array_equal(got, pf)
    np.testing.assert_array_equal(got[inner == 1], prim[inner == 1])


@pytest.mark.parametrize("precision", [0, 1])
def test_nan_in_frozen_region_does_not_stop_the_area(T, C, precision):
    mesh, prim, params, inner = _cube(C)
    cc = mesh.node_xyz[mesh.cell_nodes].mean(axis=1)
    far = int(np.argmax(cc[:, 0]))
    bad = prim.copy()
    bad[far, 0] = np.nan
    with T.Setka(mesh, dict(params, precision=precision)) as s:
        s.set_inner_area(inner)
        s.set_state(bad, 0)
        s.Go(3, is_inner_area=True)
        got = s.get_state()
    assert np.isnan(got[far, 0]) and not np.isnan(np.delete(got[:, 0], far)).any()


def test_go_area_needs_the_split(T, C):
    mesh, prim, params, _ = _cube(C)
    with T.Setka(mesh, params) as s:
        s.set_state(prim, 0)
        with pytest.raises(T.TitanError) as e:
            s.Go(1, is_inner_area=True)
        assert e.value.code == T.TITAN_E_ARG
