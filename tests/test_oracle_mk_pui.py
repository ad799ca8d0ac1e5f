"""The oracle's pickup-ion source tallies of MK_source_S (Cell::MK_Add_pui_source, Cell.cpp:1087-1133;
Cell::MK_normir_Moments, :1225-1241; Cell::MK_calc_Sm, :1135-1169), pinned by what they measure:

* S- (pui_Sm) bins the time each atom spends in the cell by its speed relative to the plasma: its
  bins sum to the density tally, and in detailed balance (atoms and protons in one Maxwellian,
  cp = 1) the histogram is the Maxwell speed distribution 4/sqrt(pi) u^2 exp(-u^2).  Binning by
  the atom's own speed (or_mk.mistake bit 3) fails it.
* S+ (pui_Sp) weighs that time by the exchange rate nu_ex: over all cells it is the expected
  exchanged weight, mean(mu) x the exchanges counted; the arise table routes each (zone, sort) to
  row k - 1 or drops it (k = 0).
* MK_calc_Sm with sigma = 1 (par_a_2 = 0) and one filled bin is the angle integral of |w - Vh|,
  ((w + Vh)^3 - |w - Vh|^3) / (3 w Vh), to the 40-point rule's accuracy; the S+ normalisation is
  per unit velocity-space volume, so S+ times the shells' volumes 4 pi/3 (w2^3 - w1^3) and the
  cell volume gives back the raw tally."""
import math

import numpy as np
import pytest

ME = 1.0 / 1836.15
A2 = 0.102046
NW, WR = 40, 6.0


def _cube(C, n=5):
    """An unjittered cube filled from every face by atoms at rest on average, protons in the same
    Maxwellian: every boundary face has the same inflow, so every particle the same weight."""
    mesh, prim, params = C.cube_brio_wu(n=n, jitter=0.0)
    mesh.face_bc = np.where(mesh.face_cells[:, 1] < 0, C.BC_OUTER_SOFT, C.BC_INTERIOR).astype(np.int8)
    prim = np.zeros((mesh.n_cells, 8))
    prim[:, 0] = 1.0
    prim[:, 4] = 1.0 / (1.0 + ME)
    return mesh, prim


def _go(O, om, mesh, prim, geom, arise, mistake=0):
    return O.mk_go(om, mesh.node_xyz, prim, 0.0, A2, 0.5, 150, seed=9, geom=geom, nthreads=4, mistake=mistake,
                   pui=dict(nw=NW, wr=WR, arise=arise))


@pytest.fixture(scope="module")
def balance(C, O):
    mesh, prim = _cube(C)
    om = O.OracleMesh(mesh)
    geom = O.measure(om, mesh.node_xyz)
    d, st, (sm, sp) = _go(O, om, mesh, prim, geom, np.ones((4, 4), np.int8))
    return mesh, prim, om, geom, d, st, sm, sp


def _maxwell_cdf(u):
    return math.erf(u) - 2.0 / math.sqrt(math.pi) * u * math.exp(-u * u)


def test_sm_bins_sum_to_the_density(balance):
    mesh, prim, om, geom, d, st, sm, sp = balance
    np.testing.assert_allclose(sm.sum(axis=1), d[:, 0] * geom["cell_volume"], rtol=1e-12)


def test_sm_histogram_is_the_maxwell_speed_distribution(C, O, balance):
    mesh, prim, om, geom, d, st, sm, sp = balance
    edges = np.arange(NW + 1) * WR / NW
    want = np.diff([_maxwell_cdf(u) for u in edges])
    want[-1] += 1.0 - _maxwell_cdf(WR)
    got = sm.sum(axis=0) / sm.sum()
    assert np.abs(got - want).max() < 0.006, np.abs(got - want).max()
    prim_moving = prim.copy()
    prim_moving[:, 1] = C.V_INF  # atoms and protons drifting together: |V| is not the relative speed
    _, _, (good, _) = O.mk_go(om, mesh.node_xyz, prim_moving, C.V_INF, A2, 0.5, 150, seed=9, geom=geom,
                              nthreads=4, pui=dict(nw=NW, wr=WR, arise=np.ones((4, 4), np.int8)))
    _, _, (bad, _) = O.mk_go(om, mesh.node_xyz, prim_moving, C.V_INF, A2, 0.5, 150, seed=9, geom=geom,
                             nthreads=4, mistake=8, pui=dict(nw=NW, wr=WR, arise=np.ones((4, 4), np.int8)))
    assert np.abs(good.sum(axis=0) / good.sum() - want).max() < 0.01
    assert np.abs(bad.sum(axis=0) / bad.sum() - want).max() > 0.05


def test_sp_is_the_exchanged_weight(C, O, balance):
    mesh, prim, om, geom, d, st, sm, sp = balance
    face, count, mu = O.mk_setup(om, geom, 0.0, 150)
    assert np.ptp(mu) < 1e-12 * mu[0] and st["births_dropped"] == 0
    assert st["exchanges"] > 15000
    assert abs(sp[:, :, 0].sum() / (mu[0] * st["exchanges"]) - 1.0) < 0.02
    assert sp[:, :, 1].sum() == 0.0
    # per bin the ratio S+ / S- is nu_ex at that relative speed: it grows with u (sigma uz)
    ratio = sp[:, :, 0].sum(axis=0)[:20] / sm.sum(axis=0)[:20]
    assert (np.diff(ratio[2:16]) > 0).all()


def test_arise_table_routes_the_rows(C, O, balance):
    """The cube's cells are all type 3: births are sort 4, exchanged atoms sort 3."""
    mesh, prim, om, geom, d, st, sm, sp = balance
    a = np.ones((4, 4), np.int8)
    a[2] = (1, 1, 0, 2)  # zone 3: sort 3 -> none, sort 4 -> row 2
    b = np.ones((4, 4), np.int8)
    b[2] = (2, 2, 1, 0)  # zone 3: sort 3 -> row 1, sort 4 -> none
    _, _, (_, spa) = _go(O, om, mesh, prim, geom, a)
    _, _, (_, spb) = _go(O, om, mesh, prim, geom, b)
    assert spa[:, :, 0].sum() == 0.0 and spa[:, :, 1].sum() > 0.0
    assert spb[:, :, 1].sum() == 0.0 and spb[:, :, 0].sum() > 0.0
    np.testing.assert_allclose(spa[:, :, 1].sum() + spb[:, :, 0].sum(), sp[:, :, 0].sum(), rtol=1e-12)


@pytest.mark.parametrize("j", [3, 11, 25])
def test_calc_sm_is_the_angle_integral(O, j):
    kn = 0.7
    sm = np.zeros((1, NW))
    sm[0, j] = 2.5
    so, _ = O.mk_pui_finish([2.0], sm, np.zeros((1, NW, 2)), WR, 0.0, kn)
    ff = 2.5 / 2.0
    vh = (j + 0.5) * WR / NW
    for ij in range(NW):
        w = (ij + 0.5) * WR / NW
        exact = ff * ((w + vh) ** 3 - abs(w - vh) ** 3) / (3.0 * w * vh) / 2.0 / kn
        assert abs(so[0, ij] / exact - 1.0) < 2e-3, (ij, so[0, ij], exact)


def test_sp_normalisation_is_per_velocity_volume(O, balance):
    mesh, prim, om, geom, d, st, sm, sp = balance
    vol = geom["cell_volume"]
    so, po = O.mk_pui_finish(vol, sm, sp, WR, A2, 0.5)
    edges = np.arange(NW + 1) * WR / NW
    shell = 4.0 * math.pi / 3.0 * (edges[1:] ** 3 - edges[:-1] ** 3)
    np.testing.assert_allclose((po * shell[None, :, None] * vol[:, None, None]).sum(axis=1), sp.sum(axis=1),
                               rtol=1e-12, atol=0)
    assert (so >= 0).all() and so.sum() > 0
