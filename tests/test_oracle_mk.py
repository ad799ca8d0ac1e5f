"""The oracle's Monte-Carlo neutrals (or_mk_go: Setka::MK_go / MK_fly_immit, Setka_MK.cpp:1225-2316)
pinned by what the kinetic theory fixes, on the cube with every boundary face an outer boundary
(bc 3) so the LISM hydrogen enters through all of them:

* free flight (no charge exchange): a convex domain filled from every face by the inflow of a
  drifting Maxwellian holds that Maxwellian, so the density is the LISM density 1 in every cell.
  A wrong inflow (Get_Spotok_inf), a birth velocity not weighted by |V.n| (Velosity_initial) or a
  birth point not uniform over the face (Get_Random_pozition) moves the cell densities off 1.
* detailed balance: with protons in the same Maxwellian (rho_Th = 1, V = (v_inf, 0, 0),
  cp = sqrt(2 p_Th / rho_Th) = 1) charge exchange leaves the distribution unchanged, so the density
  stays 1 while the exchanged atoms move to the cells' zone (sort 3).
* the exchange sampler (M_K_Change_Velosity) against the integral that defines it: mean and mean
  square of the weighted Maxwellian by quadrature; dropping its g sigma(g) weight fails it (the
  density checks above do not see that error: measured, the densities stay within noise).
* the pass through the Sun's sphere (Setka_MK.cpp:2138-2160): on a shell around the sphere the
  straight-line jump across it keeps the free-flight density at 1 in every cell past the first
  layer, the cone behind the Sun included, where absorbing the particles instead (or_mk.mistake
  bit 2) leaves a deep shadow.
* bookkeeping: the sorts sum to the total; every particle ends once (left, lost or dropped at
  birth); the particle count follows the rule of Setka_MK.cpp:1233,1380-1390."""
import math

import numpy as np
import pytest

ME = 1.0 / 1836.15


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


def _ends(st):
    return st["left"] + st["lost"]


def test_free_flight_holds_the_lism_density(C, O):
    mesh, prim, _ = _cube(C)
    om = O.OracleMesh(mesh)
    d, st = O.mk_go(om, mesh.node_xyz, prim, C.V_INF, 0.102046, 1e30, 200, seed=7)
    assert st["exchanges"] == 0 and st["lost"] == 0 and st["sun_passes"] == 0
    dropped = st["births_dropped"]
    assert _ends(st) + dropped == st["particles"]
    # MC noise of one cell ~ 0.03 at 200 particles per face; the mean over 216 cells ~ 0.002
    assert abs(d[:, 0].mean() - 1.0) < 0.012
    assert np.abs(d[:, 0] - 1.0).max() < 0.15
    np.testing.assert_allclose(d[:, 4], d[:, 0], rtol=1e-12)
    assert np.all(d[:, 1:4] == 0.0)


@pytest.mark.parametrize("kn", [1.0, 0.3])
def test_detailed_balance(C, O, kn):
    mesh, _, _ = _cube(C)
    prim = _lism(C, mesh)
    om = O.OracleMesh(mesh)
    d, st = O.mk_go(om, mesh.node_xyz, prim, C.V_INF, 0.102046, kn, 200, seed=11)
    assert st["exchanges"] > 10000
    assert abs(d[:, 0].mean() - 1.0) < 0.012
    assert np.abs(d[:, 0] - 1.0).max() < 0.15
    np.testing.assert_allclose(d[:, 1:].sum(axis=1), d[:, 0], rtol=1e-12)
    assert np.all(d[:, 1] == 0.0) and np.all(d[:, 2] == 0.0)
    assert d[:, 3].mean() > 0.15                    # exchanged atoms carry the zone's sort


def test_exchange_rate_moves_with_kn(C, O):
    """Shorter mean free path, more exchanges and a larger share of sort 3 (monotone in 1/Kn)."""
    mesh, _, _ = _cube(C, n=4)
    prim = _lism(C, mesh)
    om = O.OracleMesh(mesh)
    shares, counts = [], []
    for kn in (3.0, 1.0, 0.3):
        d, st = O.mk_go(om, mesh.node_xyz, prim, C.V_INF, 0.102046, kn, 100, seed=3)
        shares.append(d[:, 3].sum() / d[:, 0].sum())
        counts.append(st["exchanges"])
    assert counts[0] < counts[1] < counts[2]
    assert shares[0] < shares[1] < shares[2]


def test_particle_counts(C, O):
    """The count rule: mu_expect = total inflow / (N * faces); each face gets int(inflow/mu) + 1
    but at least min(N, 100) particles; faces whose inflow is below 1e-7 of the mean are skipped."""
    mesh, prim, _ = _cube(C, n=4)
    om = O.OracleMesh(mesh)
    geom = O.measure(om, mesh.node_xyz)
    face, count, mu = O.mk_setup(om, geom, C.V_INF, 50)
    b = np.flatnonzero(mesh.face_bc == C.BC_OUTER_SOFT)
    n = geom["face_normal"][b]
    ux = C.V_INF * n[:, 0]
    sp = 1.77245385
    flux = -(-np.exp(-ux * ux) + sp * ux * np.array([math.erfc(u) for u in ux])) / (2 * sp)
    flux *= geom["face_area"][b]
    keep = flux >= 1e-7 * flux.sum() / len(b)
    np.testing.assert_array_equal(face, b[keep])
    want = np.maximum((flux[keep] / (flux.sum() / (50 * len(b)))).astype(np.int64) + 1, 50)
    np.testing.assert_array_equal(count, want)
    np.testing.assert_allclose(mu * count, flux[keep], rtol=1e-13)
    # a drifting Maxwellian's outflow through a closed surface equals its inflow (the net flux
    # V . sum(n A) vanishes): the one-sided inflow formula against the outflow formula
    out = (np.exp(-ux * ux) / (2 * sp) + ux * np.array([math.erfc(-u) for u in ux]) / 2)
    np.testing.assert_allclose(flux.sum(), (out * geom["face_area"][b]).sum(), rtol=1e-9)


def _weighted_maxwellian_moments(U, V, cp, a2, npts=40):
    """Mean and mean square of the proton velocity a neutral at V meets in a charge exchange: the
    Maxwellian exp(-|w - U|^2) weighted by g sigma(g cp), g = |V - w| (units of cp), by
    Gauss-Hermite quadrature."""
    x, w = np.polynomial.hermite.hermgauss(npts)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    wt = w[:, None, None] * w[None, :, None] * w[None, None, :]
    vx, vy, vz = X + U[0], Y + U[1], Z + U[2]
    g = np.sqrt((V[0] - vx) ** 2 + (V[1] - vy) ** 2 + (V[2] - vz) ** 2)
    wt = wt * g * (1.0 - a2 * np.log(g * cp)) ** 2
    z = wt.sum()
    return np.array([(wt * vx).sum(), (wt * vy).sum(), (wt * vz).sum()]) / z, (wt * (vx ** 2 + vy ** 2 + vz ** 2)).sum() / z


@pytest.mark.parametrize("U,V,cp", [((0.0, 0.0, 0.0), (1.5, 0.3, 0.0), 1.0),
                                    ((0.4, -0.2, 0.1), (-2.0, 0.5, 0.7), 0.7)])
def test_exchange_sampler_against_its_integral(O, U, V, cp):
    """M_K_Change_Velosity draws from exp(-|w - U|^2) |V - w| sigma(|V - w| cp): 400k draws match
    the quadrature mean (MC noise ~1e-3) and mean square; the same sampler without the rejection
    step (or_mk.mistake bit 1, the g sigma(g) weight dropped) misses the mean by > 0.1."""
    m, e = _weighted_maxwellian_moments(np.array(U), np.array(V), cp, 0.102046)
    W, cap = O.mk_change_velosity(400000, U, V, cp, 0.102046, seed=3)
    assert cap == 0
    assert np.abs(W.mean(axis=0) - m).max() < 0.006
    assert abs((W ** 2).sum(axis=1).mean() - e) < 0.015
    Wb, _ = O.mk_change_velosity(400000, U, V, cp, 0.102046, seed=3, mistake=2)
    assert np.abs(Wb.mean(axis=0) - m).max() > 0.1


def test_free_flight_catches_an_unweighted_birth(C, O):
    """The birth velocity without its |V.n| weight (or_mk.mistake bit 0) leaves the free-flight
    density visibly off 1, so the free-flight pin catches that error."""
    mesh, prim, _ = _cube(C)
    om = O.OracleMesh(mesh)
    d, _ = O.mk_go(om, mesh.node_xyz, prim, C.V_INF, 0.102046, 1e30, 200, seed=7, mistake=1)
    assert abs(d[:, 0].mean() - 1.0) > 0.04


def test_threads_keep_the_histories(C, O):
    """With several OpenMP threads each particle keeps its own random stream: the counts are the
    single-thread counts and the densities agree to rounding (the tallies' order varies)."""
    mesh, _, _ = _cube(C, n=4)
    prim = _lism(C, mesh)
    om = O.OracleMesh(mesh)
    d1, s1 = O.mk_go(om, mesh.node_xyz, prim, C.V_INF, 0.102046, 1.0, 100, seed=9)
    d4, s4 = O.mk_go(om, mesh.node_xyz, prim, C.V_INF, 0.102046, 1.0, 100, seed=9, nthreads=4)
    assert s1 == s4
    np.testing.assert_allclose(d4, d1, rtol=1e-12, atol=1e-15)


def _sun_layers(O, mesh, d):
    """Mean densities of the layers of the Sun shell and of the cone behind the Sun (x < 0, within
    45 degrees of the -x axis), layers 1..4 (the first layer, whose inner hundredth the jump skips,
    is left out)."""
    g = O.measure(O.OracleMesh(mesh), mesh.node_xyz)
    cc = g["cell_center"]
    r = np.linalg.norm(cc, axis=1)
    behind = (cc[:, 0] < 0) & (np.linalg.norm(cc[:, 1:], axis=1) < -cc[:, 0])
    out = []
    for L in range(1, 5):
        m = (r > 4.0 ** (L / 8)) & (r < 4.0 ** ((L + 1) / 8))
        out.append((d[m, 0].mean(), d[m & behind, 0].mean()))
    return np.array(out)


def test_pass_through_the_sun_keeps_the_free_flight_density(C, O):
    mesh, prim, _ = C.sun_shell()
    om = O.OracleMesh(mesh)
    d, st = O.mk_go(om, mesh.node_xyz, prim, C.V_INF, 0.102046, 1e30, 1000, seed=5, nthreads=4)
    assert st["sun_passes"] > 0.03 * st["particles"] and st["lost"] == 0
    assert _ends(st) + st["births_dropped"] == st["particles"]
    lay = _sun_layers(O, mesh, d)
    # MC noise of a layer mean ~0.3 %, of the 32-cell cone ~1 %
    assert np.abs(lay[:, 0] - 1.0).max() < 0.012, lay
    assert np.abs(lay[:, 1] - 1.0).max() < 0.03, lay
    # absorbing the particles at the sphere instead leaves a shadow behind the Sun
    da, _ = O.mk_go(om, mesh.node_xyz, prim, C.V_INF, 0.102046, 1e30, 1000, seed=5, nthreads=4, mistake=4)
    assert _sun_layers(O, mesh, da)[:2, 1].max() < 0.5


@pytest.mark.parametrize("zone", [1, 2, 3, 4])
def test_thermal_proton_split(O, zone):
    """Plasma_components_1 (Phys_param.cpp:628-665) pinned by the mass it splits, rho = rho_Th
    (1 + me) + rho_He (1 + al me / 4) with al = 2 electrons per helium ion in zones 1-2 and 1 in
    3-4, by p_Th = rho_Th T_Th / 2 (T_Th is the reference's separate expression), and by the
    helium-free plasma: rho_Th = rho / (1 + me) exactly, p_Th = p / 2 to one rounding."""
    al = 2 if zone <= 2 else 1
    rng = np.random.default_rng(zone)
    for rho, p, he in zip(rng.uniform(0.1, 5.0, 50), rng.uniform(0.01, 3.0, 50), rng.uniform(0.0, 0.3, 50)):
        he = he * rho
        r_th, p_th, t_th = O.plasma_components_1(rho, p, he, zone)
        assert abs(r_th * (1.0 + ME) + he * (1.0 + al * ME / 4.0) - rho) < 1e-14 * rho
        assert abs(p_th - r_th * t_th / 2.0) < 1e-14 * p
        r0, p0, _ = O.plasma_components_1(rho, p, 0.0, zone)
        assert r0 == rho / (1.0 + ME)
        assert abs(p0 - p / 2.0) <= 2.3e-16 * p


def test_helium_takes_its_share_of_the_exchanges(C, O):
    """The flight sees the thermal protons only: with helium rho_He = 0.3 rho and p set so that
    cp = sqrt(2 p_Th / rho_Th) stays 1, detailed balance still holds (density 1) and the expected
    exchange count scales with rho_Th, here to 0.7 of the helium-free count (MC noise ~0.6%)."""
    mesh, _, _ = _cube(C, n=5)
    om = O.OracleMesh(mesh)
    prim = _lism(C, mesh)
    d0, st0 = O.mk_go(om, mesh.node_xyz, prim, C.V_INF, 0.102046, 0.3, 150, seed=21, nthreads=4)
    he = 0.3
    zones = np.asarray(mesh.cell_zone)
    al = np.where(zones <= 2, 2, 1)
    r_th = -(ME * al * he + 4.0 * (-1.0 + he)) / (4.0 * (1.0 + ME))
    ph = prim.copy()
    ph[:, 4] = (r_th / 2.0) * (8.0 + (-7.0 + al + (1.0 - al) * ME) * he) / (4.0 - (4.0 + al * ME) * he)
    scal = np.zeros((mesh.n_cells, 2))
    scal[:, 1] = he
    d1, st1 = O.mk_go(om, mesh.node_xyz, ph, C.V_INF, 0.102046, 0.3, 150, seed=22, nthreads=4, scal=scal)
    assert abs(d1[:, 0].mean() - 1.0) < 0.012
    ratio = st1["exchanges"] / st0["exchanges"]
    assert abs(ratio - r_th.mean() * (1.0 + ME)) < 0.02, ratio


def test_the_marker_routes_the_exchanged_sort(C, O):
    """P.sort = determ_zone(P.cel, 0) (Setka_MK.cpp:2088, 2121; Setka_physics.cpp:1427-1482): with
    the scalars carried, Q/rho < 50 gives sort 1 in type-1 cells and 2 elsewhere, Q/rho >= 50 gives
    4 in type-4 cells and 3 elsewhere; without them the cell's type is the sort."""
    mesh, _, _ = _cube(C, n=4)
    mesh.cell_zone = (np.arange(mesh.n_cells) % 4 + 1).astype(np.int8)  # every type
    om = O.OracleMesh(mesh)
    prim = _lism(C, mesh)
    scal = np.zeros((mesh.n_cells, 2))
    d_lo, _ = O.mk_go(om, mesh.node_xyz, prim, C.V_INF, 0.102046, 0.5, 60, seed=4, scal=scal)
    scal[:, 0] = 60.0 * prim[:, 0]
    d_hi, _ = O.mk_go(om, mesh.node_xyz, prim, C.V_INF, 0.102046, 0.5, 60, seed=4, scal=scal)
    d_no, _ = O.mk_go(om, mesh.node_xyz, prim, C.V_INF, 0.102046, 0.5, 60, seed=4)
    # columns: total, then sorts 1..4 (sort 4 is also the birth sort)
    assert d_lo[:, 1].sum() > 0 and d_lo[:, 2].sum() > 0 and d_lo[:, 3].sum() == 0
    assert d_hi[:, 1].sum() == 0 and d_hi[:, 2].sum() == 0 and d_hi[:, 3].sum() > 0
    assert (d_no[:, 1:4].sum(axis=0) > 0).all()
    # the same flights either way: only the labels move
    np.testing.assert_allclose(d_lo[:, 0], d_hi[:, 0], rtol=1e-12)
    np.testing.assert_allclose(d_lo[:, 0], d_no[:, 0], rtol=1e-12)
