"""The C ABI library loads, exports every entry point include/titan_gpu.h declares, and its
host-only parts (mesh construction, argument validation) work without a GPU."""
import ctypes
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "titan_gpu.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(titan_[a-z0-9_]+)\s*\(", src)))


def test_exports_every_declared_symbol(T):
    L = T.lib()
    names = declared_symbols()
    assert len(names) >= 25
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(T.EXPORTS)


def test_no_oracle_in_product(T):
    """The product library never links the oracle."""
    data = open(T.LIB_PATH, "rb").read()
    assert b"or_chlld" not in data and b"liboracle" not in data


def test_argument_validation_without_device(T):
    L = T.lib()
    assert L.titan_create(None, None, None) == T.TITAN_E_ARG
    assert L.titan_chlld_batch(-1, None, None, None, None, None, None, None, 0, None, None, None, None, None, None, 0) == T.TITAN_E_ARG
    bad = np.array([5], np.int8)
    q = np.zeros(8)
    n = np.array([1.0, 0, 0])
    out = np.zeros(8)
    rc = L.titan_chlld_batch(1, bad.ctypes.data, None, None, n.ctypes.data, None, q.ctypes.data, q.ctypes.data, 0,
                             None, None, out.ctypes.data, None, None, None, 0)
    assert rc == T.TITAN_E_ARG
    assert L.titan_mlp_n_params() == 9221


def test_mesh_builders(T):
    xyz, cn = T.mesh_cube(3, 0.0, 1)
    assert xyz.shape == (64, 3) and cn.shape == (27, 8)
    fn, fc, cf = T.mesh_connect(cn, 64)
    assert fn.shape[0] == 3 * 9 * 4
    # a corrupted hexahedron (repeated node) is a topology error
    cn2 = cn.copy()
    cn2[0, 1] = cn2[0, 0]
    try:
        T.mesh_connect(cn2, 64)
        raise AssertionError("expected TitanError")
    except T.TitanError as e:
        assert e.code == T.TITAN_E_TOPOLOGY
    r = np.array([1.0, 2.0, 4.0])
    xyz, cn = T.mesh_shell(3, r)
    assert xyz.shape == ((6 * 9 + 2) * 3, 3) and cn.shape == (6 * 9 * 2, 8)
    rad = np.linalg.norm(xyz, axis=1)
    assert np.allclose(np.sort(np.unique(np.round(rad, 12))), r)
