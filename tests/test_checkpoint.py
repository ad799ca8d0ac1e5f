"""The checkpoint file of Setka::Save_cell_parameters / Setka::Download_cell_parameters
(Setka_physics.cpp:4107-4285) through the library's host entry points, against the oracle's
record-by-record restatement (oracle/checkpoint.py) and a byte string assembled by hand from the
reference's write calls.  No GPU: the raw file functions are host code."""
import os
import struct

import numpy as np
import pytest

import checkpoint as CK

KEYS = ["rho", "Vx", "Vy", "Vz", "p", "Bx", "By", "Bz"]


def test_known_bytes(T, tmp_path):
    """One cell {rho: 1.5, p: 0.25} and one node (1, 2, 3): the bytes the reference's out.write
    calls produce, in this key order."""
    path = str(tmp_path / "one.bin")
    T.checkpoint_write(path, ["rho", "p"], [[1.5, 0.25]], xyz=[[1.0, 2.0, 3.0]])
    want = (struct.pack("<Q", 2) + struct.pack("<Q", 3) + b"rho" + struct.pack("<d", 1.5)
            + struct.pack("<Q", 1) + b"p" + struct.pack("<d", 0.25)
            + b"\x01" + struct.pack("<Q", 1) + struct.pack("<ddd", 1.0, 2.0, 3.0) + b"\x00" * 1000)
    assert open(path, "rb").read() == want
    vals, xyz = T.checkpoint_read(path, 1, ["p", "rho", "Q"])
    assert vals[0, 0] == 0.25 and vals[0, 1] == 1.5 and np.isnan(vals[0, 2])
    np.testing.assert_array_equal(xyz, [[1.0, 2.0, 3.0]])


def test_write_matches_oracle_reader(T, tmp_path):
    rng = np.random.default_rng(2)
    nc, nn = 257, 300
    vals = rng.normal(size=(nc, 10))
    xyz = rng.normal(size=(nn, 3))
    keys = KEYS + ["Q", "rho_H1"]
    path = str(tmp_path / "w.bin")
    T.checkpoint_write(path, keys, vals, xyz=xyz)
    cells, x = CK.read(path, nc)
    assert len(cells) == nc
    for c in range(nc):
        assert list(cells[c]) == keys
        np.testing.assert_array_equal([cells[c][k] for k in keys], vals[c])
    np.testing.assert_array_equal(np.array(x), xyz)


def test_read_any_key_order_and_extra_keys(T, tmp_path):
    """The reference writes an unordered_map: key order differs between cells, cells may hold keys
    the reader does not want, and a cell may lack a key."""
    rng = np.random.default_rng(4)
    nc = 120
    cells = []
    truth = np.full((nc, 8), np.nan)
    extra = ["Q", "rho_He", "MK_n_H", "MK_n_H1", "rho_H1", "p_H1", "Vx_H1"]
    for c in range(nc):
        names = KEYS + extra
        if c % 17 == 3:
            names = [k for k in names if k != "Bz"]
        order = list(rng.permutation(len(names))) if c % 5 else list(range(len(names)))
        kv = {}
        for i in order:
            kv[names[i]] = float(rng.normal())
        cells.append(kv)
        for j, k in enumerate(KEYS):
            if k in kv:
                truth[c, j] = kv[k]
    path = str(tmp_path / "r.bin")
    CK.write(path, cells, xyz=None)
    vals, xyz = T.checkpoint_read(path, nc, KEYS)
    assert xyz is None
    np.testing.assert_array_equal(np.isnan(vals), np.isnan(truth))
    np.testing.assert_array_equal(vals[~np.isnan(truth)], truth[~np.isnan(truth)])


def test_node_count_and_damage(T, tmp_path):
    path = str(tmp_path / "n.bin")
    T.checkpoint_write(path, ["rho"], np.ones((4, 1)), xyz=np.zeros((5, 3)))
    with pytest.raises(T.TitanError) as e:           # another mesh
        T.checkpoint_read(path, 4, ["rho"], n_nodes=6)
    assert e.value.code == T.TITAN_E_ARG
    data = open(path, "rb").read()
    for cut in (3, 20, len(data) - 1000 - 10, len(data) - 1000 - 24 * 5):
        bad = str(tmp_path / ("cut%d.bin" % cut))
        open(bad, "wb").write(data[:cut])
        with pytest.raises(T.TitanError) as e:
            T.checkpoint_read(bad, 4, ["rho"], n_nodes=5)
        assert e.value.code == T.TITAN_E_IO
    # a key length far beyond any the reference writes
    bad = str(tmp_path / "key.bin")
    open(bad, "wb").write(struct.pack("<QQ", 1, 1 << 40) + b"\x00" * 64)
    with pytest.raises(T.TitanError) as e:
        T.checkpoint_read(bad, 1, ["rho"])
    assert e.value.code == T.TITAN_E_IO
    with pytest.raises(T.TitanError) as e:
        T.checkpoint_read(str(tmp_path / "missing.bin"), 1, ["rho"])
    assert e.value.code == T.TITAN_E_IO


def test_oracle_reader_resets_mk_keys(tmp_path):
    path = str(tmp_path / "mk.bin")
    CK.write(path, [{"rho": 2.0, "MK_n_H1": 5.0}], xyz=[(0.0, 0.0, 1.0)])
    cells, xyz = CK.read(path, 1, mk_keys=("MK_n_H1",))
    assert cells[0] == {"rho": 2.0, "MK_n_H1": 0.0}
    assert xyz == [(0.0, 0.0, 1.0)]


def test_mk_parameters_known_bytes(T, tmp_path):
    """Two cells of (MK_n_H, MK_n_H1): the bytes of Save_cell_MK_parameters' out.write calls
    (Setka_physics.cpp:4173-4208)."""
    path = str(tmp_path / "mk1.bin")
    T.mk_parameters_write(path, [[0.5, 0.25], [2.0, -1.0]], keys=("MK_n_H", "MK_n_H1"))
    want = (struct.pack("<Q", 2) + struct.pack("<Q", 6) + b"MK_n_H" + struct.pack("<Q", 7) + b"MK_n_H1"
            + struct.pack("<dddd", 0.5, 0.25, 2.0, -1.0) + b"\x00" * 1000)
    assert open(path, "rb").read() == want
    got = np.full((2, 3), 9.0)
    assert T.mk_parameters_read(path, got, keys=("MK_n_H1", "Q", "MK_n_H")) == 2
    np.testing.assert_array_equal(got, [[0.25, 9.0, 0.5], [-1.0, 9.0, 2.0]])


def test_mk_parameters_against_the_oracle(T, tmp_path):
    """Writer and reader against write_mk / read_mk, with the zone being computed kept
    (Download_cell_MK_parameters' zone_except, Setka_physics.cpp:4323)."""
    rng = np.random.default_rng(7)
    nc = 301
    keys = list(T.MK_PARAMETERS)
    vals = rng.normal(size=(nc, 5))
    zone = rng.integers(1, 8, size=nc).astype(np.int8)
    path = str(tmp_path / "mk2.bin")
    T.mk_parameters_write(path, vals)
    cells = [{k: -7.0 for k in keys} for _ in range(nc)]
    assert CK.read_mk(path, cells, cell_zone=zone, zone_except=3) == keys
    mine = np.full((nc, 5), -7.0)
    assert T.mk_parameters_read(path, mine, cell_mk=zone, zone_except=3) == 5
    np.testing.assert_array_equal(mine, [[kv[k] for k in keys] for kv in cells])
    assert (mine[zone == 3] == -7.0).all() and (mine[zone != 3] == vals[zone != 3]).all()
    # the oracle's writer, the library's reader
    path2 = str(tmp_path / "mk3.bin")
    CK.write_mk(path2, keys[::-1], vals[:, ::-1])
    back = np.zeros((nc, 5))
    T.mk_parameters_read(path2, back)
    np.testing.assert_array_equal(back, vals)


def test_mk_parameters_damage(T, tmp_path):
    path = str(tmp_path / "mk4.bin")
    T.mk_parameters_write(path, np.ones((6, 5)))
    data = open(path, "rb").read()
    for cut in (4, 30, len(data) - 1000 - 8):
        bad = str(tmp_path / ("mkcut%d.bin" % cut))
        open(bad, "wb").write(data[:cut])
        with pytest.raises(T.TitanError) as e:
            T.mk_parameters_read(bad, np.zeros((6, 5)))
        assert e.value.code == T.TITAN_E_IO
    for n in (5, 7):                             # a file of another cell count
        with pytest.raises(T.TitanError) as e:
            T.mk_parameters_read(path, np.zeros((n, 5)))
        assert e.value.code == T.TITAN_E_ARG
    with pytest.raises(T.TitanError) as e:       # more cells than the file holds values for
        T.mk_parameters_read(path, np.zeros((40, 5)))
    assert e.value.code == T.TITAN_E_IO
    with pytest.raises(ValueError):
        T.mk_parameters_read(path, np.zeros((6, 4)))
    with pytest.raises(T.TitanError) as e:
        T.mk_parameters_read(str(tmp_path / "none.bin"), np.zeros((6, 5)))
    assert e.value.code == T.TITAN_E_IO


def test_pickup_source_file_known_bytes(T, tmp_path):
    """One cell's data_SpSm file (Cell::write_S_ToFile, Cell.cpp:89-125): int rows 2, cols n_w,
    size n_w, then pui_Sm and pui_Sp as Eigen stores the (2, n_w) matrix, column by column."""
    path = str(tmp_path / "func_cells_SpSm_7.bin")
    sm = np.array([1.0, 2.0, 3.0])
    sp = np.array([[10.0, 20.0], [11.0, 21.0], [12.0, 22.0]])  # [iw][k]
    T.mk_s_write(path, sm, sp)
    want = struct.pack("<iii", 2, 3, 3) + struct.pack("<3d", 1, 2, 3) + struct.pack("<6d", 10, 20, 11, 21, 12, 22)
    assert open(path, "rb").read() == want
    a, b = T.mk_s_read(path, 3)
    np.testing.assert_array_equal(a, sm)
    np.testing.assert_array_equal(b, sp)
    with pytest.raises(T.TitanError) as e:
        T.mk_s_read(path, 4)                          # other sizes
    assert e.value.code == T.TITAN_E_ARG
    bad = str(tmp_path / "cut.bin")
    open(bad, "wb").write(want[:-8])
    with pytest.raises(T.TitanError) as e:
        T.mk_s_read(bad, 3)
    assert e.value.code == T.TITAN_E_IO
