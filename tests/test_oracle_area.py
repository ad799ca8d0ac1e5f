"""Setka::Go's areas in the oracle (params area 1 / 2 with or_mesh.cell_inner; Setka_physics.cpp:113-212,
1526-1537): pins that do not restate the selection rule -- an area holding every cell is the plain
step, cells outside the area keep their state bit for bit, a face off the area's list is never
evaluated (a NaN state deep in the frozen region does not stop the step), and restricting the
faces that give dt can only lengthen the next step."""
import numpy as np


def _case(C, n=8):
    mesh, prim, params = C.cube_brio_wu(n=n)
    cc = mesh.node_xyz[mesh.cell_nodes].mean(axis=1)
    inner = (cc[:, 0] < 0.4).astype(np.uint8)
    return mesh, prim, params, inner


def test_full_area_is_the_plain_step(C, O):
    mesh, prim, params, _ = _case(C)
    om = O.OracleMesh(mesh)
    rc0, p0, _, d0, _ = O.go(om, params, mesh.node_xyz, prim, 6)
    for area, mask in ((1, np.ones(mesh.n_cells, np.uint8)), (2, np.zeros(mesh.n_cells, np.uint8))):
        om.set_inner(mask)
        rc, p, _, d, _ = O.go(om, dict(params, area=area), mesh.node_xyz, prim, 6)
        assert rc == rc0 == 0
        np.testing.assert_array_equal(p, p0)
        np.testing.assert_array_equal(d, d0)


def test_frozen_cells_keep_their_state(C, O):
    mesh, prim, params, inner = _case(C)
    om = O.OracleMesh(mesh)
    om.set_inner(inner)
    for area in (1, 2):
        rc, p, _, _, _ = O.go(om, dict(params, area=area), mesh.node_xyz, prim, 5)
        assert rc == 0
        frozen = inner == 0 if area == 1 else inner == 1
        np.testing.assert_array_equal(p[frozen], prim[frozen])
        assert np.abs(p[~frozen] - prim[~frozen]).max() > 0.0


def test_faces_off_the_list_are_not_evaluated(C, O):
    """A NaN state in a frozen cell whose faces are all off the list: the reference never touches
    it, so the inner-area step runs; the plain step stops in chlld."""
    mesh, prim, params, inner = _case(C)
    cc = mesh.node_xyz[mesh.cell_nodes].mean(axis=1)
    far = int(np.argmax(cc[:, 0]))          # x ~ 1: no neighbour is inner
    nb = set(mesh.face_cells[(mesh.face_cells == far).any(axis=1)].reshape(-1)) - {far, -1}
    assert inner[far] == 0 and all(inner[c] == 0 for c in nb)
    bad = prim.copy()
    bad[far, 0] = np.nan
    om = O.OracleMesh(mesh)
    rc_plain, *_ = O.go(om, params, mesh.node_xyz, bad, 1)
    assert rc_plain != 0
    om.set_inner(inner)
    rc, p, _, _, _ = O.go(om, dict(params, area=1), mesh.node_xyz, bad, 3)
    assert rc == 0 and np.isnan(p[far, 0]) and not np.isnan(np.delete(p[:, 0], far)).any()


def test_area_dt_from_its_faces_only(C, O):
    """The min over a subset of the faces is never below the min over all of them; here the
    fastest face lies in the outer area, so the inner area's second step is strictly longer."""
    mesh, prim, params, inner = _case(C)
    om = O.OracleMesh(mesh)
    _, _, _, d_all, _ = O.go(om, params, mesh.node_xyz, prim, 2)
    om.set_inner(inner)
    _, _, _, d_in, _ = O.go(om, dict(params, area=1), mesh.node_xyz, prim, 2)
    _, _, _, d_out, _ = O.go(om, dict(params, area=2), mesh.node_xyz, prim, 2)
    assert d_in[0] == d_out[0] == d_all[0] == 1e-5
    assert d_in[1] >= d_all[1] and d_out[1] >= d_all[1]
    assert min(d_in[1], d_out[1]) == d_all[1]
    assert max(d_in[1], d_out[1]) > d_all[1]
