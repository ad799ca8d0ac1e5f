"""Parity of the passive scalars Q and rho_He with the oracle: their face fluxes (the konvect part of
Setka::Culc_Gran_Potok, Setka_physics.cpp:2349-2361, 2485-2493), their update and clips in whole
steps (:1781-1817, 2004-2016) in both precisions, on areas and with the fused surrogate, the
closed-form source's zone routing by Q (determ_zone, :1427-1482), and checkpoints.  Size-independent
pins on the device: scalars equal to rho follow rho bit for bit, scalars of 2 rho and rho / 4 stay in
that ratio (the oracle's tests give the reasons)."""
import importlib

import numpy as np
import pytest

import checkpoint as CK
from helpers import cell_err, cell_errors

pytestmark = pytest.mark.gpu


def _scal(prim, seed=1):
    """Q / rho spread over [1, 100] (across determ_zone's switch at 50), rho_He in [0.05, 0.3] rho
    and above rho in every 37th cell (clipped by the first update)."""
    rng = np.random.default_rng(seed)
    n = prim.shape[0]
    q = prim[:, 0] * (1.0 + 99.0 * rng.random(n))
    he = prim[:, 0] * (0.05 + 0.25 * rng.random(n))
    he[::37] = 1.5 * prim[::37, 0]
    return np.column_stack([q, he])


def _shell(C, moving=False, tvd=False, n_edge=6, n_layers=10):
    mesh, prim, params = C.shell(n_edge=n_edge, n_layers=n_layers, moving=moving, r_zone1_au=80.0, noise=0.01)
    return mesh, prim, dict(params, tvd=1 if tvd else 0, passive_scalars=1)


def _rel(a, b):
    """Largest relative difference, per value, with the scale of each column as the floor."""
    scale = np.maximum(np.abs(b).max(axis=0, keepdims=True) * 1e-300, np.abs(b))
    scale = np.where(scale > 0, scale, 1.0)
    return float((np.abs(a - b) / scale).max())


def _gpu_go(T, mesh, prim, scal, params, steps, h=None):
    with T.Setka(mesh, params) as s:
        s.set_state(prim, 0)
        s.set_scalars(scal, 0)
        if h is not None:
            s.set_h_fluids(h)
        st = s.Go(steps)
        return s.get_state(), s.get_scalars(), st


@pytest.mark.parametrize("tvd", [False, True])
def test_face_fluxes_cube_bit_exact(T, C, O, tvd):
    mesh, prim, params = C.cube_brio_wu(n=12)
    params = dict(params, tvd=1 if tvd else 0, passive_scalars=1)
    scal = _scal(prim)
    with T.Setka(mesh, params) as s:
        s.set_state(prim, 0)
        s.set_scalars(scal, 0)
        flux, dtmin, br, kflux = s.Culc_Gran_Potok(0, 1e-5, want_scalars=True)
        g = s.geometry(0)
    rc, rflux, rdt, rbr, rk = O.face_flux(O.OracleMesh(mesh, tvd=tvd), params, prim, g, scal=scal)
    assert rc == 0
    np.testing.assert_array_equal(br, rbr)
    np.testing.assert_array_equal(flux, rflux)
    np.testing.assert_array_equal(kflux, rk)


def _upwind_scale(mesh, prim, scal, area, gamma):
    """Per face and scalar, a bound of every term of its konvect flux: the larger cell value of the
    scalar x the larger |V| + fast speed of the two cells x the area."""
    a = mesh.face_cells[:, 0]
    b = np.where(mesh.face_cells[:, 1] >= 0, mesh.face_cells[:, 1], a)
    rho, p, v, B = prim[:, 0], prim[:, 4], prim[:, 1:4], prim[:, 5:8]
    cf = np.sqrt(gamma * p / rho + (B * B).sum(axis=1) / (4.0 * np.pi * rho))
    sp = np.linalg.norm(v, axis=1) + cf
    q = np.maximum(np.abs(scal[a]), np.abs(scal[b]))
    return q * np.maximum(sp[a], sp[b])[:, None] * area[:, None]


def test_face_fluxes_shell_zone1_tvd(T, C, O):
    """Faces that reach a Zone_1 cell use the transcendental frame functions (last-ulp differences
    between the device and glibc): there the error is bounded by 1e-12 of the flux's upwind terms
    (q (|V| + c_f) A; the flux itself can be far smaller on faces along the wind, where v.n nearly
    vanishes), bit for bit elsewhere."""
    mesh, prim, params = _shell(C, tvd=True)
    scal = C.passive_scalars(mesh, prim)
    with T.Setka(mesh, params) as s:
        s.set_state(prim, 0)
        s.set_scalars(scal, 0)
        flux, _, br, kflux = s.Culc_Gran_Potok(0, 1e-5, want_scalars=True)
        g = s.geometry(0)
    om = O.OracleMesh(mesh, tvd=True)
    tvd = om.arrs["face_tvd"]
    rc, rflux, _, rbr, rk = O.face_flux(om, params, prim, g, scal=scal)
    assert rc == 0
    np.testing.assert_array_equal(br, rbr)
    cells = np.column_stack([mesh.face_cells, tvd])
    z1 = np.zeros(mesh.n_faces, bool)
    for k in range(4):
        c = cells[:, k]
        z1 |= (c >= 0) & (mesh.cell_zone[np.maximum(c, 0)] == 1)
    assert z1.any() and (~z1).any()
    np.testing.assert_array_equal(kflux[~z1], rk[~z1])
    scale = _upwind_scale(mesh, prim, scal, g["face_area"], params["gamma"])[z1]
    err = np.abs(kflux[z1] - rk[z1]) / scale
    print("Zone_1 faces: scalar flux error / upwind scale %.2e, bit-exact faces %d of %d"
          % (err.max(), int((kflux == rk).all(axis=1).sum()), mesh.n_faces))
    assert err.max() < 1e-12


@pytest.mark.parametrize("tvd", [False, True])
def test_go_cube_bit_exact(T, C, O, tvd):
    mesh, prim, params = C.cube_brio_wu(n=12)
    params = dict(params, tvd=1 if tvd else 0, passive_scalars=1)
    scal = _scal(prim)
    got, gsc, _ = _gpu_go(T, mesh, prim, scal, params, 40)
    rc, ref, _, _, _, rsc = O.go(O.OracleMesh(mesh, tvd=tvd), params, mesh.node_xyz, prim, 40, nthreads=8, scal=scal)
    assert rc == 0
    np.testing.assert_array_equal(got, ref)
    np.testing.assert_array_equal(gsc, rsc)


@pytest.mark.parametrize("moving", [False, True])
@pytest.mark.parametrize("tvd", [False, True])
def test_go_shell_vs_oracle(T, C, O, moving, tvd):
    mesh, prim, params = _shell(C, moving=moving, tvd=tvd)
    scal = C.passive_scalars(mesh, prim)
    got, gsc, _ = _gpu_go(T, mesh, prim, scal, params, 30)
    rc, ref, _, _, _, rsc = O.go(O.OracleMesh(mesh, tvd=tvd), params, mesh.node_xyz, prim, 30, nthreads=8,
                                 disp=mesh.node_disp, scal=scal)
    assert rc == 0
    e = [_rel(gsc[:, k], rsc[:, k]) for k in range(2)]
    print("shell moving=%s tvd=%s, 30 steps: per-cell errors" % (moving, tvd), cell_errors(got, ref),
          "Q %.2e rho_He %.2e" % tuple(e))
    assert cell_err(got, ref) < 1e-12
    assert max(e) < 1e-12


def test_scalars_leave_the_mhd_state_alone(T, C):
    mesh, prim, params = _shell(C, moving=True, tvd=True)
    got, _, _ = _gpu_go(T, mesh, prim, C.passive_scalars(mesh, prim), params, 20)
    with T.Setka(mesh, dict(params, passive_scalars=0)) as s:
        s.set_state(prim, 0)
        s.Go(20)
        ref = s.get_state()
    np.testing.assert_array_equal(got, ref)


def test_identity_and_ratio_on_device(T, C):
    """FP64 on a larger moving shell (17280 cells), 20 steps: scalars equal to rho stay equal to it
    bit for bit (with and without TVD); without TVD, 2 rho and rho / 4 stay exactly in that ratio."""
    for tvd in (True, False):
        mesh, prim, params = _shell(C, moving=True, tvd=tvd, n_edge=12, n_layers=20)
        rb = mesh.face_bstate[:, 0]
        mesh.face_bscalar = np.column_stack([rb, rb])
        got, sc, _ = _gpu_go(T, mesh, prim, np.column_stack([prim[:, 0], prim[:, 0]]), params, 20)
        np.testing.assert_array_equal(sc[:, 0], got[:, 0])
        np.testing.assert_array_equal(sc[:, 1], got[:, 0])
        if not tvd:
            mesh.face_bscalar = np.column_stack([2.0 * rb, 0.25 * rb])
            got, sc, _ = _gpu_go(T, mesh, prim, np.column_stack([2.0 * prim[:, 0], 0.25 * prim[:, 0]]), params, 20)
            np.testing.assert_array_equal(sc[:, 0], 2.0 * got[:, 0])
            np.testing.assert_array_equal(sc[:, 1], 0.25 * got[:, 0])


def test_mixed_precision_scalars(T, C, O):
    """The mixed-precision step: its scalar flux is homogeneous in the scalar (Q = 8 rho_He stays
    exactly so, without TVD), and after 30 steps of the moving shell the scalars agree with the FP64
    oracle per cell to 1e-6 (the bound of the mixed step's conserved variables)."""
    mesh, prim, params = _shell(C, moving=True)
    params = dict(params, precision=1)
    rng = np.random.default_rng(6)
    base = prim[:, 0] * (0.1 + 0.8 * rng.random(mesh.n_cells))
    rb = mesh.face_bstate[:, 0] * 0.5
    mesh.face_bscalar = np.column_stack([2.0 * rb, 0.25 * rb])
    _, sc, st = _gpu_go(T, mesh, prim, np.column_stack([2.0 * base, 0.25 * base]), params, 30)
    np.testing.assert_array_equal(sc[:, 0], 8.0 * sc[:, 1])
    scal = C.passive_scalars(mesh, prim)
    mesh.face_bscalar = C._hard_face_scalars(mesh)
    _, gsc, _ = _gpu_go(T, mesh, prim, scal, params, 30)
    rc, _, _, _, _, rsc = O.go(O.OracleMesh(mesh), dict(params, precision=0), mesh.node_xyz, prim, 30, nthreads=8,
                               disp=mesh.node_disp, scal=scal)
    assert rc == 0
    e = [float((np.abs(gsc[:, k] - rsc[:, k]) / np.abs(rsc[:, k])).max()) for k in range(2)]
    print("mixed, 30 steps: per-cell scalar errors Q %.2e rho_He %.2e" % tuple(e))
    assert max(e) < 1e-6


@pytest.mark.parametrize("is_inner", [True, False])
def test_go_area_with_scalars_bit_exact(T, C, O, is_inner):
    mesh, prim, params = C.cube_brio_wu(n=10, jitter=0.1)
    params = dict(params, passive_scalars=1)
    cc = mesh.node_xyz[mesh.cell_nodes].mean(axis=1)
    inner = (cc[:, 0] < 0.4).astype(np.uint8)
    scal = _scal(prim, seed=2)
    with T.Setka(mesh, params) as s:
        s.set_inner_area(inner)
        s.set_state(prim, 0)
        s.set_scalars(scal, 0)
        s.Go(7, is_inner_area=is_inner)
        got, gsc = s.get_state(), s.get_scalars()
    om = O.OracleMesh(mesh)
    om.set_inner(inner)
    rc, pf, _, _, _, rsc = O.go(om, dict(params, area=1 if is_inner else 2), mesh.node_xyz, prim, 7, scal=scal)
    assert rc == 0
    np.testing.assert_array_equal(got, pf)
    np.testing.assert_array_equal(gsc, rsc)
    frozen = (inner == 0) if is_inner else (inner == 1)
    np.testing.assert_array_equal(gsc[frozen], scal[frozen])


def _banded(C):
    """The shell with Zone_1..4 in radial bands of three layers, and Q / rho across 50."""
    mesh, prim, params = _shell(C, n_edge=4, n_layers=12)
    r_in = np.linalg.norm(mesh.node_xyz, axis=1)[mesh.cell_nodes].min(axis=1)
    layer = np.searchsorted(np.unique(np.round(r_in, 9)), np.round(r_in, 9))
    mesh.cell_zone = (1 + np.minimum(layer // 3, 3)).astype(np.int8)
    return mesh, prim, params


def test_bera_sources_follow_the_zone(T, C, O):
    """titan_sources(BERA) with the scalars: determ_zone's routing (Zone_4 cells with Q/rho >= 50 take
    al = 1) and rho_He in the proton split, against the oracle at 1e-12 relative to each variable's
    scale (the device's log and the host's differ in the last ulp)."""
    mesh, prim, params = _banded(C)
    scal = _scal(prim, seed=3)
    h = C.h_fluids(mesh.n_cells)
    with T.Setka(mesh, dict(params, source_mode=T.SRC_BERA)) as s:
        s.set_state(prim, 0)
        s.set_scalars(scal, 0)
        s.set_h_fluids(h)
        got = s.sources(0, T.SRC_BERA)
    ref = O.bera_source(prim, h, scal=scal, cell_type=mesh.cell_zone)
    zone = np.array([O.determ_zone(scal[c, 0], prim[c, 0], mesh.cell_zone[c]) for c in range(mesh.n_cells)])
    assert set(zone.tolist()) == {1, 2, 3, 4}
    scale = np.abs(ref).max(axis=0)
    err = np.abs(got - ref).max(axis=0) / np.maximum(scale, 1e-300)
    assert err.max() < 1e-12, err
    # the routing matters: al = 1 cells differ from the al = 2 split
    al2 = O.bera_source(prim, h, scal=scal, cell_type=np.minimum(mesh.cell_zone, 3))
    assert np.abs(al2 - ref)[zone == 4].max() > 0.0


def test_go_bera_with_scalars(T, C, O):
    mesh, prim, params = _shell(C)
    scal = C.passive_scalars(mesh, prim)
    h = C.h_fluids(mesh.n_cells)
    p = dict(params, source_mode=T.SRC_BERA)
    got, gsc, _ = _gpu_go(T, mesh, prim, scal, p, 20, h=h)
    rc, ref, _, _, _, rsc = O.go(O.OracleMesh(mesh), p, mesh.node_xyz, prim, 20, nthreads=8, h_fluids=h, scal=scal)
    assert rc == 0
    print("closed-form source with scalars, 20 steps: per-cell errors", cell_errors(got, ref))
    assert cell_err(got, ref) < 1e-12
    assert max(_rel(gsc[:, k], rsc[:, k]) for k in range(2)) < 1e-12


def test_fused_surrogate_step_scalars(T, C, O):
    """The surrogate step updates the scalars in its own phase, fallback cells included; they do not
    depend on the source, so one step agrees with the oracle's step without a source."""
    sur = importlib.import_module("titan_c-_amd.surrogate")
    mesh, prim, params = _shell(C, moving=True)
    nh = C.neutrals(mesh.n_cells)
    x = sur.features(prim[:, 0], prim[:, 4] / prim[:, 0], prim[:, 1:4], nh[:, 0], nh[:, 1], nh[:, 2:5])
    norm = sur.Normaliser(x, np.ones((len(x), 5)), box_quantile=0.1)  # some cells fall back
    in_norm, _, box = norm.blocks()
    scal = C.passive_scalars(mesh, prim)
    with T.Setka(mesh, dict(params, source_mode=T.SRC_MLP, cx_nodes=32)) as s:
        s.set_state(prim, 0)
        s.set_scalars(scal, 0)
        s.set_neutrals(nh)
        s.mlp_set(sur.init_params(0), in_norm, np.full(5, 0.1, np.float32), box)
        st = s.Go(1)
        gsc = s.get_scalars()
    assert st["n_fallback"] > 0
    rc, _, _, _, _, rsc = O.go(O.OracleMesh(mesh), params, mesh.node_xyz, prim, 1, disp=mesh.node_disp, scal=scal)
    assert rc == 0
    assert max(_rel(gsc[:, k], rsc[:, k]) for k in range(2)) < 1e-12


def test_checkpoint_round_trip_with_scalars(T, C, tmp_path):
    mesh, prim, params = _shell(C, n_edge=4, n_layers=6)
    scal = _scal(prim, seed=4)
    path = str(tmp_path / "q.bin")
    with T.Setka(mesh, params) as s:
        s.set_state(prim, 0)
        s.set_scalars(scal, 0)
        s.Go(3)
        state, sc = s.get_state(), s.get_scalars()
        s.Save_cell_parameters(path)
    cells, _ = CK.read(path, mesh.n_cells)
    np.testing.assert_array_equal(np.array([[c["Q"], c["rho_He"]] for c in cells]), sc)
    with T.Setka(mesh, params) as s2:
        s2.Download_cell_parameters(path)
        np.testing.assert_array_equal(s2.get_state(0), state)
        np.testing.assert_array_equal(s2.get_scalars(0), sc)
        np.testing.assert_array_equal(s2.get_scalars(1), sc)


def test_scalar_calls_need_the_switch(T, C):
    mesh, prim, params = C.cube_brio_wu(n=4)
    with T.Setka(mesh, params) as s:
        with pytest.raises(T.TitanError):
            s.set_scalars(np.zeros((mesh.n_cells, 2)), 0)
        s.set_state(prim, 0)
        with pytest.raises(T.TitanError):
            s.Culc_Gran_Potok(0, 1e-5, want_scalars=True)


@pytest.mark.parametrize("table", [False, True])
def test_bera_fluid_sources_vs_oracle(T, C, O, table):
    """titan_sources_h: the hydrogen fluids' sources in every zone (default routing, or a caller's
    arise table), per cell and fluid, against the oracle at 1e-12 of each variable's per-cell scale
    (the fluid's loss term), and the plasma part equal to titan_sources(BERA)."""
    mesh, prim, params = _banded(C)
    scal = _scal(prim, seed=3)
    h = C.h_fluids(mesh.n_cells)
    ar = np.array([[2, 1, 4, 3], [1, 1, 1, 1], [4, 4, 2, 2], [3, 4, 1, 2]], np.int8) if table else None
    with T.Setka(mesh, dict(params, source_mode=T.SRC_BERA)) as s:
        s.set_state(prim, 0)
        s.set_scalars(scal, 0)
        s.set_h_fluids(h)
        plasma, got = s.sources_h(0, 4, ar)
        plain = s.sources(0, T.SRC_BERA)
    np.testing.assert_array_equal(plasma, plain)
    ref_p, ref = O.bera_source_h(prim, h, scal=scal, cell_type=mesh.cell_zone, arise=ar)
    np.testing.assert_array_equal(ref_p, O.bera_source(prim, h, scal=scal, cell_type=mesh.cell_zone))
    scale = np.abs(ref).max(axis=1, keepdims=True) + 1e-300  # per cell and variable, over the fluids
    err = (np.abs(got - ref) / scale).max()
    print(f"fluid sources table={table}: max per-cell error {err:.3e}", flush=True)
    assert err < 1e-12
    assert np.abs(ref[:, :, 1:]).max() > 0.0
