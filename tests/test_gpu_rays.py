"""The B, C, D, E, H and G rays on the device (titan_set_rays_ext, titan_dvigenie, titan_go with
move_mesh 2) against the oracle's Luch::dvigenie restatement (Luch.cpp:153-739), bit for bit: the
device takes each ray's cos and sin of its angles from the host, as the oracle computes them."""
import math

import numpy as np
import pytest

from helpers import EXT_M, ext_ray_set
from test_gpu_fit import _perturbed, _setup

pytestmark = pytest.mark.gpu


def _a_rays(rng, first, n_rays, g, xyz):
    """A rays on nodes first.. with their TS, HP and BS nodes placed along the ray."""
    M0, M1, M11, M2, M3, M4 = (g[k] for k in ("M0", "M1", "M11", "M2", "M3", "M4"))
    npr = M0 + M1 + M2 + M3 + M4 + 5
    nodes = (first + np.arange(n_rays * npr)).reshape(n_rays, npr).astype(np.int32)
    ang = np.stack([rng.uniform(0.1, 1.2, n_rays), rng.uniform(0, 2 * math.pi, n_rays)], axis=1)
    for k in range(n_rays):
        d = np.array([math.cos(ang[k, 0]), math.sin(ang[k, 0]) * math.cos(ang[k, 1]),
                      math.sin(ang[k, 0]) * math.sin(ang[k, 1])])
        for j, r in ((M0 + M1 + 1, 80.0), (M0 + M1 + M2 + 2, 120.0), (M0 + M1 + M2 + M3 + 3, 240.0)):
            xyz[nodes[k, j]] = r * (1.0 + 0.05 * rng.uniform(-1, 1)) * d
    return dict(ray_nodes=nodes, ray_angles=ang, R0=g["R0"], R1=g["R1"], R5=g["R5"], M0=M0, M1=M1, M11=M11, M2=M2,
                M3=M3, M4=M4, da1=1.3, da2=0.8, da3=0.7, da4=1.2, da5=0.5, phi_a=math.pi / 3)


@pytest.mark.parametrize("counts,serial,n_cube", [((6,) * 6, (), 14), ((40,) * 6, (5, 6), 20),
                                                  ((300,) * 6, (), 30)])
def test_dvigenie_ext_matches_oracle(T, C, O, counts, serial, n_cube):
    mesh, _, params = C.cube_brio_wu(n=n_cube)
    rng = np.random.default_rng(sum(counts) + n_cube)
    n_a = 4
    npr = sum(EXT_M[k] for k in ("M0", "M1", "M2", "M3", "M4")) + 5
    types, lists, opor, ang, g, xyz = ext_ray_set(rng, n_a * npr, mesh.n_nodes, counts, serial)
    ra = _a_rays(rng, 0, n_a, g, xyz)
    print(f"ext rays counts={counts} serial={serial} nodes={mesh.n_nodes}", flush=True)
    with T.Setka(mesh, params) as s:
        s.set_rays(ra)
        s.set_rays_ext(types, lists, opor, ang, g)
        s.set_nodes(xyz, 0)
        s.dvigenie(0)
        got = s.get_nodes(0)
    F = O.OracleFit(mesh, ra)
    want = O.dvigenie_ext(O.RaysExt(types, lists, opor, ang, g), O.dvigenie(F, xyz))
    assert np.abs(want - xyz).max() > 1.0
    np.testing.assert_array_equal(got, want)


def test_set_rays_ext_rejects_bad_sets(T, C):
    mesh, _, params = C.cube_brio_wu(n=10)
    rng = np.random.default_rng(1)
    npr = sum(EXT_M[k] for k in ("M0", "M1", "M2", "M3", "M4")) + 5
    types, lists, opor, ang, g, xyz = ext_ray_set(rng, npr, mesh.n_nodes, (2, 2, 2, 2, 2, 2))
    ra = _a_rays(rng, 0, 1, g, xyz)
    with T.Setka(mesh, params) as s:
        with pytest.raises(T.TitanError):  # before titan_set_rays
            s.set_rays_ext(types, lists, opor, ang, g)
        s.set_rays(ra)
        bad = [l[:] for l in lists]
        bad[0] = bad[0][:-1]  # a B ray one node short
        with pytest.raises(T.TitanError):
            s.set_rays_ext(types, bad, opor, ang, g)
        op = opor.copy()
        op[2 * 2, 0] = -1  # a D ray without its C-ray node
        with pytest.raises(T.TitanError):
            s.set_rays_ext(types, lists, op, ang, g)
        with pytest.raises(T.TitanError):
            s.set_rays_ext(types, lists, opor, ang, dict(g, N5=1))
        s.set_rays_ext(types, lists, opor, ang, g)


def _b_shell(C, O, zone1, cos_lim=0.3):
    """The fitted shell with its flank rays (|cos the| < cos_lim) run as B rays (support nodes their
    R1, TS, HP and BS nodes) and started where Luch::dvigenie puts them."""
    mesh, prim, params, rays = _perturbed(C, zone1, seed=7)
    ang = rays["ray_angles"]
    flank = np.abs(np.cos(ang[:, 0])) < cos_lim
    M0, M1, M2, M3 = rays["M0"], rays["M1"], rays["M2"], rays["M3"]
    rb = rays["ray_nodes"][flank]
    ra = dict(rays, ray_nodes=rays["ray_nodes"][~flank], ray_angles=ang[~flank])
    opor = np.stack([rb[:, M0], rb[:, M0 + M1 + 1], rb[:, M0 + M1 + M2 + 2], rb[:, M0 + M1 + M2 + M3 + 3]], axis=1)
    geo = dict(O.EXT_GEO_DEFAULTS, N4=0, N5=2, MF=0)
    ext = O.RaysExt(np.full(len(rb), 1, np.int8), list(rb), opor, ang[flank], dict(geo, **{
        k: rays[k] for k in ("R0", "R1", "R5", "M0", "M1", "M11", "M2", "M3", "M4")}))
    F = O.OracleFit(mesh, ra)
    mesh.node_xyz = O.dvigenie_ext(ext, O.dvigenie(F, mesh.node_xyz))
    return mesh, prim, params, ra, ext, geo


@pytest.mark.parametrize("zone1", [False, True])
def test_go_fitted_with_b_rays_matches_oracle(T, C, O, zone1):
    """titan_go with move_mesh 2 on a shell whose flank rays are B rays: each step moves the
    surfaces, the A rays, then the B rays, before the flux."""
    mesh, prim, params, ra, ext, geo = _b_shell(C, O, zone1)
    print(f"b-ray shell zone1={zone1} b_rays={len(ext.type)} a_rays={len(ra['ray_nodes'])}", flush=True)
    s, F = _setup(T, O, mesh, params, ra)
    F.ext = ext
    steps = 6
    with s:
        s.set_rays_ext(ext.type, [ext.nodes[ext.off[k]:ext.off[k + 1]] for k in range(len(ext.type))], ext.opor,
                       ext.angles, geo)
        s.set_state(prim, 0)
        s.Go(steps)
        got = s.get_state()
        gx = s.get_nodes(s.current_layer())
    om = O.OracleMesh(mesh)
    rc, pf, xf, _, _ = O.go(om, params, mesh.node_xyz, prim, steps, fit=F)
    assert rc == 0
    hp = ext.opor[:, 2]
    assert np.abs(xf[hp] - mesh.node_xyz[hp]).max() > 0.0
    if zone1:
        np.testing.assert_allclose(gx, xf, rtol=1e-12, atol=1e-12 * np.abs(xf).max())
        scale = np.abs(pf).max(axis=0) + 1e-300
        assert (np.abs(got - pf) / scale).max() < 1e-9
    else:
        np.testing.assert_array_equal(gx, xf)
        np.testing.assert_array_equal(got, pf)
