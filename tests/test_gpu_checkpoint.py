"""Checkpoints of a device context: Setka::Save_cell_parameters / Download_cell_parameters
(Setka_physics.cpp:4107-4285) through titan_save_cell_parameters / titan_download_cell_parameters,
and Setka::Init_TVD again after new nodes (titan_init_tvd)."""
import numpy as np
import pytest

import checkpoint as CK

pytestmark = pytest.mark.gpu

KEYS = ["rho", "Vx", "Vy", "Vz", "p", "Bx", "By", "Bz"]


def _moving_cube(C, n=10):
    mesh, prim, params = C.cube_brio_wu(n=n)
    rng = np.random.default_rng(3)
    xyz = mesh.node_xyz
    inner = np.all((xyz > 1e-9) & (xyz < 1 - 1e-9), axis=1)
    disp = np.zeros_like(xyz)
    disp[inner] = 0.01 * rng.normal(size=(inner.sum(), 3))
    mesh.node_disp = disp
    return mesh, prim, dict(params, move_mesh=1, motion_omega=3.0, motion_phase=0.4)


def test_save_writes_newest_layer(T, C, tmp_path):
    mesh, prim, params = _moving_cube(C)
    q = np.arange(mesh.n_cells, dtype=np.float64) * 0.5
    path = str(tmp_path / "s.bin")
    with T.Setka(mesh, params) as s:
        s.set_state(prim, 0)
        s.Go(7)
        state = s.get_state()
        nodes = s.get_nodes(s.current_layer())
        s.Save_cell_parameters(path, extra={"Q": q})
    cells, xyz = CK.read(path, mesh.n_cells)
    got = np.array([[c[k] for k in KEYS] for c in cells])
    np.testing.assert_array_equal(got, state)
    np.testing.assert_array_equal([c["Q"] for c in cells], q)
    np.testing.assert_array_equal(np.array(xyz), nodes)
    assert np.abs(nodes - mesh.node_xyz).max() > 0.0


def test_download_equals_set_state(T, C, tmp_path):
    """A downloaded state steps exactly like the same state handed over by titan_set_state; keys a
    cell lacks in the file keep the context's values."""
    mesh, prim, params = C.cube_brio_wu(n=10)
    rng = np.random.default_rng(8)
    saved = prim * (1.0 + 0.01 * rng.random(prim.shape))
    cells = []
    for c in range(mesh.n_cells):
        kv = {k: saved[c, j] for j, k in enumerate(KEYS)}
        if c % 11 == 0:
            del kv["By"]                       # kept from the context
        kv["MK_n_H1"] = 3.0                    # not a plasma key: skipped
        cells.append(dict(reversed(list(kv.items()))))
    path = str(tmp_path / "d.bin")
    CK.write(path, cells, xyz=None)
    expect = saved.copy()
    expect[::11, 6] = prim[::11, 6]
    with T.Setka(mesh, params) as s:
        s.set_state(prim, 0)
        s.set_state(prim, 1)
        s.Download_cell_parameters(path)
        np.testing.assert_array_equal(s.get_state(0), expect)
        np.testing.assert_array_equal(s.get_state(1), expect)
        s.Go(5)
        a = s.get_state()
    with T.Setka(mesh, params) as s:
        s.set_state(expect, 0)
        s.Go(5)
        b = s.get_state()
    np.testing.assert_array_equal(a, b)


def test_download_nodes_remeasures_and_init_tvd(T, C, O, tmp_path):
    """New nodes from a checkpoint replace both node layers; the next titan_go measures the geometry
    of the new nodes (as Setka::Go does on entry), and titan_init_tvd picks the TVD neighbours the
    oracle's Init_TVD picks for them."""
    mesh, prim, params = C.cube_brio_wu(n=10, jitter=0.1)
    rng = np.random.default_rng(12)
    xyz1 = mesh.node_xyz.copy()
    inner = np.all((xyz1 > 1e-9) & (xyz1 < 1 - 1e-9), axis=1)
    xyz1[inner] += 0.02 * rng.normal(size=(inner.sum(), 3))
    path = str(tmp_path / "n.bin")
    T.checkpoint_write(path, KEYS, prim, xyz=xyz1)
    p = dict(params, tvd=1)
    with T.Setka(mesh, p) as s:
        s.set_state(prim, 0)
        s.Go(2)                                  # started: geometry of the old nodes
        s.Download_cell_parameters(path)
        np.testing.assert_array_equal(s.get_nodes(0), xyz1)
        np.testing.assert_array_equal(s.get_nodes(1), xyz1)
        s.Go(1)
        g = s.geometry(s.current_layer())
        s.Init_TVD()
        tvd = s.tvd_cells()
    om = O.OracleMesh(mesh)
    ref = O.measure(om, xyz1)
    for k in ("cell_center", "cell_volume", "face_center", "face_normal", "face_area"):
        np.testing.assert_array_equal(g[k], ref[k], err_msg=k)
    np.testing.assert_array_equal(tvd, O.init_tvd(om, ref))
    assert (tvd != O.OracleMesh(mesh, tvd=True).arrs["face_tvd"]).any()
