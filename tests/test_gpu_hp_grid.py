"""HP smoothing on the ray grids and the corrections after the move (titan_smooth_angle_hp,
titan_set_hp_grid) on the device against the oracle's Setka::Culc_Velocity_surface restatement
(Setka_move.cpp:700-1357), with the set-up order of TITAN.cpp:108-112.  The azimuthal macros1
terms take polar angles through atan, which the device and glibc round differently in the last
bit, so velocities and nodes agree to 1e-11 relative rather than bit for bit."""
import numpy as np
import pytest

from helpers import hp_grid_lists
from test_gpu_fit import _perturbed

pytestmark = pytest.mark.gpu

N4 = 6


def _setup(T, O, C, zone1=False, phi_a=1.0):
    mesh, prim, params, rays = _perturbed(C, zone1)
    rays = dict(rays, phi_a=phi_a)
    x0 = mesh.node_xyz
    L = hp_grid_lists(rays, x0)
    s = T.Setka(mesh, params)
    s.set_rays(rays)
    F = O.OracleFit(mesh, rays)
    # TITAN.cpp:108-112
    s.Find_Yzel_Sosed_for_BS()
    s.Smooth_angle_HP(L["angle_nodes"], L["nb6"])
    s.Smooth_head_HP3()
    s.Smooth_head_TS3()
    F.smooth(3, x0)
    cent = [T.circle_fit(x0[[int(v)] + [int(u) for u in nb]][:, 1:3])[0] for v, nb in zip(L["angle_nodes"], L["nb6"])]
    F.smooth_angle_hp(L["angle_nodes"], L["nb6"], x0, centres=cent)
    F.smooth(2, x0)
    F.smooth(1, x0)
    # the sphere centres the library fitted (its fit and the oracle's agree to 1e-9, not bitwise)
    phi_lim = F.phi_lim
    for i, v in enumerate(F.sn):
        a = x0[v]
        near = O.polar_angle(a[0], float(np.hypot(a[1], a[2]))) <= phi_lim
        if F.sn_kind[i] == 1 or (F.sn_kind[i] == 2 and near):
            if len(F.cand[i]) >= 3:
                F.centre[i] = T.sphere_fit(x0[F.cand[i]])
    grid = dict(tail=L["tail"], N4=N4, bs_ref=L["bs_ref"], bs_lists=L["bs_lists"])
    s.set_hp_grid(L["grid_nodes"], L["nb8"], **grid)
    F.set_hp_grid(L["grid_nodes"], L["nb8"], **grid)
    return s, F, mesh, prim, params, L


def test_circle_fit_matches_oracle(T, O):
    rng = np.random.default_rng(7)
    for _ in range(20):
        c, R = rng.uniform(-1, 1, 2), rng.uniform(20, 40)
        a = rng.uniform(0, 6.3) + rng.uniform(0.1, 0.9) * np.arange(7)
        P = c + R * np.stack([np.cos(a), np.sin(a)], axis=1) + rng.normal(0, 0.05, (7, 2))
        cl, rl = T.circle_fit(P)
        co, ro = O.circle_fit(P)
        np.testing.assert_allclose(cl, co, rtol=0, atol=1e-9 * R)
        assert abs(rl - ro) < 1e-9 * R


@pytest.mark.parametrize("zone1", [False, True])
def test_culc_velocity_surface_with_hp_grid_matches_oracle(T, C, O, zone1):
    s, F, mesh, prim, params, L = _setup(T, O, C, zone1)
    print(f"hp grid zone1={zone1} grid={len(L['grid_nodes'])} angle={len(L['angle_nodes'])} "
          f"tails={len(L['tail'])} bs={len(L['bs_ref'])}", flush=True)
    x0 = mesh.node_xyz
    dt = 3.0e-3
    with s:
        s.set_state(prim, 0)
        v = s.Culc_Velocity_surface(0, dt)
        x1 = s.get_nodes(1)
    om = O.OracleMesh(mesh)
    rc, xn, vel = O.culc_velocity_surface(om, params, F, x0, prim, dt)
    assert rc == 0
    moved = F.sn
    np.testing.assert_allclose(v[moved], vel, rtol=1e-11, atol=1e-11 * np.abs(vel).max())
    np.testing.assert_allclose(x1[moved], xn[moved], rtol=1e-13, atol=1e-13 * np.abs(x0).max())
    # the grid terms act: without them the grid nodes' velocities differ
    G = O.OracleFit(mesh, dict(F.rays, phi_a=1.0))
    G.nb, G.centre = F.nb, F.centre
    rc, _, vel0 = O.culc_velocity_surface(om, params, G, x0, prim, dt)
    gi = [F._sn_index[int(u)] for u in L["grid_nodes"]]
    assert np.abs(vel[gi] - vel0[gi]).max() > 1e-6 * np.abs(vel).max()


def test_go_fitted_with_hp_grid_matches_oracle(T, C, O):
    s, F, mesh, prim, params, L = _setup(T, O, C, False)
    steps = 6
    with s:
        s.set_state(prim, 0)
        s.Go(steps)
        got = s.get_state()
        gx = s.get_nodes(s.current_layer())
    om = O.OracleMesh(mesh)
    rc, pf, xf, _, _ = O.go(om, params, mesh.node_xyz, prim, steps, fit=F)
    assert rc == 0
    np.testing.assert_allclose(gx, xf, rtol=1e-11, atol=1e-11 * np.abs(xf).max())
    scale = np.abs(pf).max(axis=0) + 1e-300
    assert (np.abs(got - pf) / scale).max() < 1e-9


def test_repeated_abscissa_reports_surface_error(T, C, O):
    """solveQuadraticEquation stops the reference on two equal abscissae (Help_funk.cpp:293-298):
    the library returns TITAN_E_SURFACE naming the node."""
    s, F, mesh, prim, params, L = _setup(T, O, C, False)
    nb8 = L["nb8"].copy()
    nb8[3, 1] = nb8[3, 0]  # AA11 = AA1
    with s:
        s.set_hp_grid(L["grid_nodes"], nb8)
        s.set_state(prim, 0)
        with pytest.raises(T.TitanError) as e:
            s.Culc_Velocity_surface(0, 3.0e-3)
    assert e.value.code == -8 and e.value.index == int(L["grid_nodes"][3])
    F.set_hp_grid(L["grid_nodes"], nb8)
    rc, _, _ = O.culc_velocity_surface(O.OracleMesh(mesh), params, F, mesh.node_xyz, prim, 3.0e-3)
    assert rc == -9
