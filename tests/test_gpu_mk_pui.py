"""The pickup-ion source tallies on the device (titan_set_mk_source_s, titan_mk_get_s,
titan_mk_s_finish_batch: Cell::MK_Add_pui_source, MK_normir_Moments, MK_calc_Sm, Cell.cpp:1087-1241)
against the oracle.  The finishing pass is deterministic and matches to 1e-12 (the device log); the
tallies come from flights that share the oracle's random streams, so they are compared within MC
noise, and through the identity the oracle is pinned by: S+ integrated over velocity space and the
cells is the exchanged weight."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ME = 1.0 / 1836.15
A2 = 0.102046
NW, WR, KN = 40, 6.0, 0.5


def _cube(C, n=5):
    mesh, _, params = C.cube_brio_wu(n=n, jitter=0.0)
    mesh.face_bc = np.where(mesh.face_cells[:, 1] < 0, C.BC_OUTER_SOFT, C.BC_INTERIOR).astype(np.int8)
    prim = np.zeros((mesh.n_cells, 8))
    prim[:, 0] = 1.0
    prim[:, 4] = 1.0 / (1.0 + ME)
    return mesh, prim, dict(params, par_a_2=A2, par_Kn=KN, v_inf=0.0)


def test_finish_matches_the_oracle(T, O):
    rng = np.random.default_rng(3)
    n = 60
    sm = rng.uniform(0.0, 1.0, (n, NW)) * (rng.uniform(size=(n, NW)) > 0.3)
    sm[5] = 0.0                                    # an empty cell
    sp = rng.uniform(0.0, 1.0, (n, NW, 2))
    vol = rng.uniform(0.5, 2.0, n)
    so, po = T.mk_s_finish_batch(vol, sm, sp, WR, A2, KN)
    oso, opo = O.mk_pui_finish(vol, sm, sp, WR, A2, KN)
    np.testing.assert_allclose(so, oso, rtol=1e-12, atol=0)
    np.testing.assert_array_equal(po, opo)
    assert (so[5] == 0).all()


def test_tallies_match_the_oracle(T, C, O):
    mesh, prim, params = _cube(C)
    ones = np.ones((4, 4), np.int8)
    om = O.OracleMesh(mesh)
    geom = O.measure(om, mesh.node_xyz)
    vol = geom["cell_volume"]
    with T.Setka(mesh, params) as s:
        s.set_state(prim, 0)
        s.MK_source_S(NW, WR, ones)
        d, st = s.MK_go(150, seed=9, layer=0)
        sm, sp = s.mk_S()
    od, ost, (rsm, rsp) = O.mk_go(om, mesh.node_xyz, prim, 0.0, A2, KN, 150, seed=9, geom=geom, nthreads=4,
                                  pui=dict(nw=NW, wr=WR, arise=ones))
    osm, osp = O.mk_pui_finish(vol, rsm, rsp, WR, A2, KN)
    assert st["particles"] == ost["particles"]
    # the same random stream per particle: totals far inside the MC noise
    for a, b in ((sm, osm), (sp[:, :, 0], osp[:, :, 0])):
        ta, tb = (a * vol[:, None]).sum(axis=0), (b * vol[:, None]).sum(axis=0)
        assert abs(ta.sum() / tb.sum() - 1.0) < 0.01
        big = tb > 0.02 * tb.max()
        assert np.abs(ta[big] / tb[big] - 1.0).max() < 0.05
    assert (sp[:, :, 1] == 0).all()
    # S+ over velocity space and the cells: the exchanged weight
    edges = np.arange(NW + 1) * WR / NW
    shell = 4.0 * math.pi / 3.0 * (edges[1:] ** 3 - edges[:-1] ** 3)
    face, count, mu = O.mk_setup(om, geom, 0.0, 150)
    total = (sp[:, :, 0] * shell[None, :] * vol[:, None]).sum()
    assert abs(total / (mu[0] * st["exchanges"]) - 1.0) < 0.02


def test_reference_tables_and_arguments(T, C):
    mesh, prim, params = _cube(C, n=4)
    with T.Setka(mesh, params) as s:
        s.set_state(prim, 0)
        with pytest.raises(T.TitanError):
            s.mk_S()                                    # no tallies yet
        with pytest.raises(T.TitanError):
            s.MK_source_S(NW, WR, np.full((4, 4), 3))    # no fluid 3
        with pytest.raises(T.TitanError):
            s.MK_source_S(5000, WR)
        s.MK_source_S(NW, WR)                           # proton_arise of the run with pickup ions
        s.MK_go(60, seed=2, layer=0)
        sm, sp = s.mk_S()
        # type-3 cells: births (sort 4) feed fluid 1, the exchanged atoms (sort 3) nothing
        assert sm.sum() > 0 and sp[:, :, 0].sum() > 0 and (sp[:, :, 1] == 0).all()
        s.MK_source_S(0, 0.0)
        s.MK_go(10, seed=2, layer=0)
        with pytest.raises(T.TitanError):
            s.mk_S()
