"""Pins the oracle's mesh indexing and geometry (Setka::New_connect, Setka.cpp:2066-2249;
Gran::Culc_measure, Gran.cpp:112-218; Cell::Culc_volume, Cell.cpp:958-1000; Gran::culc_velosity,
Gran.cpp:571-589) by exact values on axis-aligned boxes, the invariants of Setka::Test_geometr
(Setka.cpp:475-617), closure and conservation identities, and two independent implementations of
the face numbering (oracle C and the library's C++)."""
import numpy as np
import pytest


def cube_mesh(C, n, jitter=0.0, seed=1):
    mesh, prim, params = C.cube_brio_wu(n=n, jitter=jitter, seed_mesh=seed)
    return mesh


def test_new_connect_two_cells(O):
    """Brute force on two hexahedra sharing the face x = 1: 11 faces; the shared face is created by
    cell 0 as its local face 4 (nodes 1,2,6,5), cell 1 is its cells[1], and each cell lists its
    faces in ascending id."""
    xs = [0, 1, 2]
    nodes = {}
    for k in range(2):
        for j in range(2):
            for i in range(3):
                nodes[(i, j, k)] = len(nodes)
    def hexa(i):
        return [nodes[(i, 0, 0)], nodes[(i + 1, 0, 0)], nodes[(i + 1, 1, 0)], nodes[(i, 1, 0)],
                nodes[(i, 0, 1)], nodes[(i + 1, 0, 1)], nodes[(i + 1, 1, 1)], nodes[(i, 1, 1)]]
    cn = np.array([hexa(0), hexa(1)], np.int32)
    fn, fc, cf = O.new_connect(cn, len(nodes))
    assert fn.shape[0] == 11
    c0 = cn[0]
    local = [(0, 3, 2, 1), (4, 5, 6, 7), (0, 1, 5, 4), (2, 3, 7, 6), (1, 2, 6, 5), (0, 4, 7, 3)]
    for l in range(6):
        np.testing.assert_array_equal(fn[l], c0[list(local[l])])
    assert tuple(fc[4]) == (0, 1)
    assert all(fc[f, 1] == -1 for f in range(11) if f != 4)
    np.testing.assert_array_equal(cf[0], np.arange(6))
    assert 4 in cf[1] and np.all(np.diff(cf[1]) > 0)


def test_new_connect_matches_library(T, C, O):
    for mesh in (cube_mesh(C, 7, 0.15), C.shell(n_edge=5, n_layers=4)[0]):
        fn, fc, cf = O.new_connect(mesh.cell_nodes, mesh.n_nodes)
        np.testing.assert_array_equal(fn, mesh.face_nodes)
        np.testing.assert_array_equal(fc, mesh.face_cells)
        np.testing.assert_array_equal(cf, mesh.cell_faces)


def test_face_counts(C):
    n = 9
    m = cube_mesh(C, n)
    assert m.n_faces == 3 * n * n * (n + 1)
    ne, nl = 5, 4
    sh = C.shell(n_edge=ne, n_layers=nl)[0]
    assert sh.n_cells == 6 * ne * ne * nl
    assert sh.n_nodes == (6 * ne * ne + 2) * (nl + 1)
    assert sh.n_faces == 6 * ne * ne * (nl + 1) + 12 * ne * ne * nl


def test_axis_aligned_exact(C, O):
    n = 4
    m = cube_mesh(C, n)
    g = O.measure(O.OracleMesh(m), m.node_xyz)
    h = 1.0 / n
    assert np.abs(g["face_area"] - h * h).max() < 1e-16
    assert np.abs(g["cell_volume"] - h ** 3).max() < 1e-17
    nrm = g["face_normal"]
    assert np.all(np.sort(np.abs(nrm), axis=1)[:, :2] == 0.0)
    assert np.all(np.abs(nrm).max(axis=1) == 1.0)
    cc = g["cell_center"]
    np.testing.assert_allclose(np.sort(np.unique(np.round(cc[:, 0], 12))), (np.arange(n) + 0.5) * h)


@pytest.mark.parametrize("gcl", [False, True])
def test_test_geometr_invariants(C, O, gcl):
    for m in (cube_mesh(C, 8, 0.15), C.shell(n_edge=5, n_layers=4)[0]):
        g = O.measure(O.OracleMesh(m), m.node_xyz, gcl=gcl)
        nrm, fc, cc = g["face_normal"], g["face_center"], g["cell_center"]
        assert np.abs(np.linalg.norm(nrm, axis=1) - 1).max() < 1e-14          # Test 5
        a, b = m.face_cells[:, 0], m.face_cells[:, 1]
        inner = b >= 0
        d = np.einsum("ij,ij->i", nrm[inner], cc[b[inner]] - cc[a[inner]])
        assert np.all(d > 0)                                                    # Test 1
        assert np.all(np.einsum("ij,ij->i", nrm, fc - cc[a]) > 0)
        assert all(len(set(r)) == 4 for r in m.face_nodes[:200])                # Test 6
        assert m.cell_faces.shape[1] == 6 and m.cell_nodes.shape[1] == 8        # Test 2
        assert np.all(g["cell_volume"] > 0) and np.all(g["face_area"] > 0)


@pytest.mark.parametrize("gcl", [False, True])
def test_unit_cube_volume_and_translation(C, O, gcl):
    m = cube_mesh(C, 10, 0.15)
    om = O.OracleMesh(m)
    g = O.measure(om, m.node_xyz, gcl=gcl)
    assert abs(g["cell_volume"].sum() - 1.0) < 1e-12
    g2 = O.measure(om, m.node_xyz + np.array([0.3, -1.7, 2.2]), gcl=gcl)
    assert np.abs(g2["cell_volume"] / g["cell_volume"] - 1).max() < 1e-11
    assert np.abs(g2["face_area"] / g["face_area"] - 1).max() < 1e-11


def test_gcl_closure_and_planar_agreement(C, O):
    m = cube_mesh(C, 8, 0.15)
    om = O.OracleMesh(m)
    g = O.measure(om, m.node_xyz, gcl=True)
    S = g["face_normal"] * g["face_area"][:, None]
    sign = np.where(m.face_cells[m.cell_faces, 0] == np.arange(m.n_cells)[:, None], 1.0, -1.0)
    closure = (S[m.cell_faces] * sign[:, :, None]).sum(axis=1)
    assert np.abs(closure).max() < 1e-16
    # an affine image of an axis-aligned box has planar faces: both formulas agree
    A = np.array([[1.0, 0.2, 0.1], [0.0, 0.9, -0.3], [0.1, 0.0, 1.2]])
    m2 = cube_mesh(C, 5)
    x = m2.node_xyz @ A.T
    om2 = O.OracleMesh(m2)
    gr, gg = O.measure(om2, x), O.measure(om2, x, gcl=True)
    assert np.abs(gr["cell_volume"] / gg["cell_volume"] - 1).max() < 1e-12
    assert abs(gr["cell_volume"].sum() - np.linalg.det(A)) < 1e-12
    assert np.abs(gr["face_normal"] - gg["face_normal"]).max() < 1e-12


def test_face_velocity_rigid_translation(C, O):
    m = cube_mesh(C, 6, 0.15)
    om = O.OracleMesh(m)
    g = O.measure(om, m.node_xyz)
    d = np.array([1e-3, -2e-3, 5e-4])
    dt = 0.01
    for gcl in (False, True):
        gg = O.measure(om, m.node_xyz, gcl=gcl)
        w = O.face_velocities(om, m.node_xyz, m.node_xyz + d, gg, dt, gcl=gcl)
        exp = gg["face_normal"] @ d / dt
        assert np.abs(w - exp).max() < 1e-9 * np.abs(d).max() / dt


def test_gcl_swept_volumes_match_volume_change(C, O):
    m = cube_mesh(C, 6, 0.15)
    om = O.OracleMesh(m)
    rng = np.random.default_rng(2)
    x1 = m.node_xyz
    x2 = x1 + 0.01 * rng.normal(size=x1.shape) / 6
    dt = 0.1
    g1 = O.measure(om, x1, gcl=True)
    g2 = O.measure(om, x2, gcl=True)
    w = O.face_velocities(om, x1, x2, g1, dt, gcl=True)
    sign = np.where(m.face_cells[m.cell_faces, 0] == np.arange(m.n_cells)[:, None], 1.0, -1.0)
    swept = (w * g1["face_area"] * dt)[m.cell_faces]
    dv = (swept * sign).sum(axis=1)
    assert np.abs(dv - (g2["cell_volume"] - g1["cell_volume"])).max() < 1e-15


def test_det_cos(O):
    xs = np.concatenate([np.linspace(-50, 50, 2001), [0.0, np.pi, 1e3 + 0.1, -777.7]])
    got = np.array([O.det_cos(x) for x in xs])
    assert np.abs(got - np.cos(xs)).max() < 4e-16 * 1e3
    assert np.abs(got[np.abs(xs) < 60] - np.cos(xs[np.abs(xs) < 60])).max() < 1e-15
