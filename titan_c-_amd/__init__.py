"""titan_c-_amd: MI355X (gfx950) implementation of TITAN's explicit ideal-MHD step.

Python view of the C ABI in include/titan_gpu.h (libtitan_gpu.so, built by build.py).  The class
`Setka` mirrors the reference's step interface (TITAN/source/Setka.h:121,205,221,223):
``Setka.Go``, ``Setka.Culc_Gran_Potok``, ``Setka.Calculating_measure``; ``chlld`` mirrors
``Phys_param::chlld`` (Phys_param.h:240-251) for a batch of faces.  All compute runs in the HIP
library; this module only marshals arrays.  There is no CPU fallback: loading fails loudly when
the library is missing.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libtitan_gpu.so")

TITAN_OK = 0
TITAN_E_ARG = -1
TITAN_E_HIP = -2
TITAN_E_RIEMANN = -3
TITAN_E_NONFINITE_STATE = -4
TITAN_E_TOPOLOGY = -5
TITAN_E_NOT_TRAINED = -6
TITAN_E_IO = -7
TITAN_E_SURFACE = -8
BC_INTERIOR, BC_INNER_HARD, BC_OUTER_HARD, BC_OUTER_SOFT, BC_ZERO_GRADIENT = 0, 1, 2, 3, 4
GEOM_REF, GEOM_GCL = 0, 1
PREC_FP64, PREC_MIXED = 0, 1
SRC_NONE, SRC_EXACT, SRC_MLP, SRC_BERA = 0, 1, 2, 3

_ERRNAMES = {
    TITAN_E_ARG: "TITAN_E_ARG", TITAN_E_HIP: "TITAN_E_HIP", TITAN_E_RIEMANN: "TITAN_E_RIEMANN",
    TITAN_E_NONFINITE_STATE: "TITAN_E_NONFINITE_STATE", TITAN_E_TOPOLOGY: "TITAN_E_TOPOLOGY",
    TITAN_E_NOT_TRAINED: "TITAN_E_NOT_TRAINED", TITAN_E_IO: "TITAN_E_IO", TITAN_E_SURFACE: "TITAN_E_SURFACE",
}


# titan_mk_go's counters
MK_STATS = ("particles", "exchanges", "left", "sun_passes", "lost", "births_dropped", "segments", "exchange_capped")
# titan_mk_go_zone's: the same, then the flights that ended on the zone's border
MK_ZONE_STATS = MK_STATS + ("zone_border",)
# the order in which Setka::Algoritm(2) runs the Monte-Carlo zones (Setka.cpp:200-226)
MK_ZONE_SEQUENCE = (6, 4, 2, 1, 3, 5, 7, 5, 3, 1, 2, 4, 6, 7, 4, 5, 2, 3, 1)


class TitanError(RuntimeError):
    def __init__(self, code, what, index=-1):
        self.code = code
        self.index = index
        super().__init__("%s returned %s (%d), index %d" % (what, _ERRNAMES.get(code, "?"), code, index))


class MeshC(ctypes.Structure):
    _fields_ = [("n_nodes", ctypes.c_int64), ("n_cells", ctypes.c_int64), ("n_faces", ctypes.c_int64),
                ("cell_nodes", ctypes.c_void_p), ("face_nodes", ctypes.c_void_p), ("face_cells", ctypes.c_void_p),
                ("cell_faces", ctypes.c_void_p), ("face_bc", ctypes.c_void_p), ("face_method", ctypes.c_void_p),
                ("face_flags", ctypes.c_void_p), ("face_bstate", ctypes.c_void_p), ("cell_zone", ctypes.c_void_p),
                ("cell_is_tvd", ctypes.c_void_p), ("node_xyz", ctypes.c_void_p), ("node_disp", ctypes.c_void_p),
                ("face_bscalar", ctypes.c_void_p)]


class ParamsC(ctypes.Structure):
    _fields_ = [("gamma", ctypes.c_double), ("kfl", ctypes.c_double), ("v_inf", ctypes.c_double),
                ("move_mesh", ctypes.c_int32), ("n_disc", ctypes.c_int32), ("geom_mode", ctypes.c_int32),
                ("precision", ctypes.c_int32), ("source_mode", ctypes.c_int32), ("cx_nodes", ctypes.c_int32),
                ("motion_omega", ctypes.c_double), ("motion_phase", ctypes.c_double),
                ("tvd", ctypes.c_int32), ("snos_on_hp", ctypes.c_int32),
                ("par_a_2", ctypes.c_double), ("par_n_H_LISM", ctypes.c_double), ("par_Kn", ctypes.c_double),
                ("move_TS", ctypes.c_int32), ("move_HP", ctypes.c_int32), ("move_BS", ctypes.c_int32),
                ("sglag_TS", ctypes.c_int32), ("sglag_HP", ctypes.c_int32), ("sglag_BS", ctypes.c_int32),
                ("velocity_TS", ctypes.c_double), ("velocity_HP", ctypes.c_double), ("velocity_BS", ctypes.c_double),
                ("sglag_TS_k_sphere", ctypes.c_double), ("sglag_HP_k_sphere", ctypes.c_double),
                ("sglag_BS_k", ctypes.c_double), ("passive_scalars", ctypes.c_int32)]

# parameters a params dict may lack: Phys_param.h:99-101 for par_*, Phys_param.cpp:500-520 for the
# fitted surfaces
_PARAM_DEFAULTS = {"tvd": 0, "snos_on_hp": 0, "par_a_2": 0.102046, "par_n_H_LISM": 3.0, "par_Kn": 40.0906,
                   "move_TS": 1, "move_HP": 1, "move_BS": 1, "sglag_TS": 1, "sglag_HP": 1, "sglag_BS": 0,
                   "velocity_TS": 0.1, "velocity_HP": 0.1, "velocity_BS": 0.3, "sglag_TS_k_sphere": 0.01,
                   "sglag_HP_k_sphere": 0.01, "sglag_BS_k": 0.05, "passive_scalars": 0}


class RaysC(ctypes.Structure):
    _fields_ = [("n_rays", ctypes.c_int64), ("nodes_per_ray", ctypes.c_int32), ("ray_nodes", ctypes.c_void_p),
                ("ray_angles", ctypes.c_void_p), ("R0", ctypes.c_double), ("R1", ctypes.c_double),
                ("R5", ctypes.c_double), ("M0", ctypes.c_int32), ("M1", ctypes.c_int32), ("M11", ctypes.c_int32),
                ("M2", ctypes.c_int32), ("M3", ctypes.c_int32), ("M4", ctypes.c_int32), ("da1", ctypes.c_double),
                ("da2", ctypes.c_double), ("da3", ctypes.c_double), ("da4", ctypes.c_double), ("da5", ctypes.c_double),
                ("phi_a", ctypes.c_double)]


class RaysExtC(ctypes.Structure):
    _fields_ = [("n_rays", ctypes.c_int64), ("ray_type", ctypes.c_void_p), ("ray_off", ctypes.c_void_p),
                ("ray_nodes", ctypes.c_void_p), ("ray_opor", ctypes.c_void_p), ("ray_angles", ctypes.c_void_p),
                ("L6", ctypes.c_double), ("L7", ctypes.c_double), ("N4", ctypes.c_int32), ("N5", ctypes.c_int32),
                ("MF", ctypes.c_int32)] + \
               [(k, ctypes.c_double) for k in ("dd1", "dd2", "dd3", "dd4", "dd5", "dd6", "dd7", "dd8", "ba1", "ba2",
                                               "ba3", "ba4", "ba5", "md1", "md2", "md3", "ea1", "ea2", "ea3")]


class HpGridC(ctypes.Structure):
    _fields_ = [("n_grid", ctypes.c_int64), ("grid_node", ctypes.c_void_p), ("grid_nb", ctypes.c_void_p),
                ("sglag_HP_k_angle", ctypes.c_double), ("sglag_HP_k", ctypes.c_double),
                ("sglag_HP_angle", ctypes.c_double), ("sglag_HP_along", ctypes.c_double),
                ("n_tail", ctypes.c_int32), ("tail_len", ctypes.c_int32), ("N4", ctypes.c_int32),
                ("tail", ctypes.c_void_p), ("n_bs", ctypes.c_int32), ("bs_ref", ctypes.c_void_p),
                ("bs_off", ctypes.c_void_p), ("bs_nodes", ctypes.c_void_p)]


# Phys_param.cpp:513-519
HP_GRID_DEFAULTS = dict(sglag_HP_k_angle=0.001, sglag_HP_k=0.03, sglag_HP_angle=1.8, sglag_HP_along=1.0)


# Luch::type of the other ray families (Luch.cpp:153-739)
RAY_B, RAY_C, RAY_D, RAY_E, RAY_H, RAY_G = 1, 2, 3, 4, 5, 6


class StatsC(ctypes.Structure):
    _fields_ = [("steps_done", ctypes.c_int64), ("dt_last", ctypes.c_double), ("dt_next", ctypes.c_double),
                ("t_sim", ctypes.c_double), ("n_rho_floor", ctypes.c_int64), ("n_p_floor", ctypes.c_int64),
                ("n_fallback", ctypes.c_int64), ("ms_total", ctypes.c_double),
                ("n_fp64_faces", ctypes.c_int64)]


_lib = None
_P = ctypes.c_void_p
_I = ctypes.c_int
_I64 = ctypes.c_int64
_I32 = ctypes.c_int32
_D = ctypes.c_double

_SIGS = {
    "titan_create": (_I, [ctypes.POINTER(_P), ctypes.POINTER(MeshC), ctypes.POINTER(ParamsC)]),
    "titan_destroy": (_I, [_P]),
    "titan_last_error": (_I, [_P, ctypes.POINTER(_I64)]),
    "titan_set_state": (_I, [_P, _P, _I]),
    "titan_get_state": (_I, [_P, _P, _I]),
    "titan_set_nodes": (_I, [_P, _P, _I]),
    "titan_get_nodes": (_I, [_P, _P, _I]),
    "titan_set_neutrals": (_I, [_P, _P]),
    "titan_set_h_fluids": (_I, [_P, _I, _P]),
    "titan_calculating_measure": (_I, [_P, _I]),
    "titan_get_geometry": (_I, [_P, _I, _P, _P, _P, _P, _P]),
    "titan_set_scalars": (_I, [_P, _P, _I]),
    "titan_get_scalars": (_I, [_P, _P, _I]),
    "titan_gran_potok": (_I, [_P, _I, _D, _P, _P, _P, _P]),
    "titan_go": (_I, [_P, _I64, ctypes.POINTER(StatsC)]),
    "titan_go_area": (_I, [_P, _I, _I64, ctypes.POINTER(StatsC)]),
    "titan_set_inner_area": (_I, [_P, _P]),
    "titan_current_layer": (_I, [_P]),
    "titan_tvd_cells": (_I, [_P, _P]),
    "titan_chlld_batch": (_I, [_I64, _P, _P, _P, _P, _P, _P, _P, _I, _P, _P, _P, _P, _P, _P, _I]),
    "titan_chlld_batch_noalloc": (_I, [_I64, _P, _P, _P, _P, _P, _P, _P, _I, _P, _P, _P, _P, _P, _P, _I, _P]),
    "titan_sources": (_I, [_P, _I, _I, _P]),
    "titan_sources_h": (_I, [_P, _I, _P, _P, _P]),
    "titan_mk_go": (_I, [_P, _I, ctypes.c_int32, ctypes.c_uint64, ctypes.c_int32, _P, _P]),
    "titan_set_mk_zone": (_I, [_P, ctypes.c_int32, _P, _P]),
    "titan_mk_go_zone": (_I, [_P, _I, ctypes.c_int32, ctypes.c_int32, ctypes.c_uint64, ctypes.c_int32, _P, _P]),
    "titan_mk_dist_get": (_I, [_P, _I64, ctypes.c_int32, ctypes.c_int32, _P]),
    "titan_mk_dist_set": (_I, [_P, _I64, ctypes.c_int32, ctypes.c_int32, _P]),
    "titan_set_mk_source_s": (_I, [_P, ctypes.c_int32, _D, _P]),
    "titan_mk_get_s": (_I, [_P, _P, _P]),
    "titan_mk_s_finish_batch": (_I, [_I64, ctypes.c_int32, _D, _D, _D, _P, _P, _P, _P, _P]),
    "titan_mk_s_write": (_I, [ctypes.c_char_p, ctypes.c_int32, _P, _P]),
    "titan_mk_s_read": (_I, [ctypes.c_char_p, ctypes.c_int32, _P, _P]),
    "titan_cx_exact_batch": (_I, [_I64, _P, _I, _P]),
    "titan_mlp_n_params": (_I64, []),
    "titan_mlp_set": (_I, [_P, _P, _P, _P, _P]),
    "titan_mlp_get": (_I, [_P, _P]),
    "titan_mlp_forward": (_I, [_P, _I64, _P, _P]),
    "titan_mlp_loss_grad": (_I, [_P, _I64, _P, _P, _P, _P]),
    "titan_mlp_train": (_I, [_P, _I64, _P, _P, _I, _I, ctypes.c_float, ctypes.c_uint64, _P]),
    "titan_checkpoint_write": (_I, [ctypes.c_char_p, _I64, _I, _P, _P, _I64, _P]),
    "titan_checkpoint_read": (_I, [ctypes.c_char_p, _I64, _I, _P, _P, _I64, _P, ctypes.POINTER(_I64)]),
    "titan_mk_parameters_write": (_I, [ctypes.c_char_p, _I64, _I, _P, _P]),
    "titan_mk_parameters_read": (_I, [ctypes.c_char_p, _I64, _I, _P, _P, _P, _I32, ctypes.POINTER(_I32)]),
    "titan_save_cell_parameters": (_I, [_P, ctypes.c_char_p, _I, _P, _P]),
    "titan_download_cell_parameters": (_I, [_P, ctypes.c_char_p]),
    "titan_init_tvd": (_I, [_P]),
    "titan_set_rays": (_I, [_P, ctypes.POINTER(RaysC)]),
    "titan_set_rays_ext": (_I, [_P, ctypes.POINTER(RaysExtC)]),
    "titan_smooth_angle_hp": (_I, [_P, _I64, _P, _P]),
    "titan_set_hp_grid": (_I, [_P, ctypes.POINTER(HpGridC)]),
    "titan_circle_fit": (_I, [_I64, _P, _P]),
    "titan_smooth_head_ts3": (_I, [_P]),
    "titan_smooth_head_hp3": (_I, [_P]),
    "titan_find_yzel_sosed_for_bs": (_I, [_P]),
    "titan_culc_velocity_surface": (_I, [_P, _I, _D, _P]),
    "titan_dvigenie": (_I, [_P, _I]),
    "titan_sphere_fit": (_I, [_I64, _P, _P]),
    "titan_mesh_connect": (_I64, [_I64, _P, _I64, _P, _P, _P]),
    "titan_mesh_cube": (_I, [ctypes.c_int32, _D, ctypes.c_uint64, _P, _P]),
    "titan_mesh_shell": (_I, [ctypes.c_int32, ctypes.c_int32, _P, _P, _P]),
    "titan_kernel_ms": (_D, [_P, _I]),
    "titan_set_timing": (_I, [_P, _I]),
    "titan_tile_info": (_I, [_P, _P]),
}
EXPORTS = tuple(_SIGS)


def lib():
    """Loads libtitan_gpu.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libtitan_gpu.so is not built (%s); run titan_c-_amd/build.py" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data


def _c(a, dtype):
    return None if a is None else np.ascontiguousarray(a, dtype=dtype)


def _check(rc, what, ctx=None):
    if rc != TITAN_OK:
        idx = -1
        if ctx is not None:
            i = ctypes.c_int64(-1)
            lib().titan_last_error(ctx, ctypes.byref(i))
            idx = i.value
        raise TitanError(rc, what, idx)


# ---- checkpoint files ---------------------------------------------------------------------------
def _keys(keys):
    arr = (ctypes.c_char_p * len(keys))(*[k.encode() for k in keys])
    return arr, ctypes.cast(arr, ctypes.c_void_p)


def checkpoint_write(path, keys, values, xyz=None):
    """Setka::Save_cell_parameters file: values [n_cells, len(keys)]; xyz [n_nodes, 3] or None."""
    vals = _c(values, np.float64).reshape(-1, len(keys))
    arr, kp = _keys(keys)
    x = None if xyz is None else _c(xyz, np.float64).reshape(-1, 3)
    _check(lib().titan_checkpoint_write(os.fsencode(path), vals.shape[0], len(keys), kp, _ptr(vals),
                                        -1 if x is None else x.shape[0], _ptr(x)), "titan_checkpoint_write")


def checkpoint_read(path, n_cells, keys, n_nodes=-1):
    """Setka::Download_cell_parameters file: returns (values [n_cells, len(keys)] with NaN where a
    cell lacks a key, xyz [n_nodes, 3] or None if the file has no node block).  n_nodes < 0 reads
    the node count first."""
    arr, kp = _keys(keys)
    vals = np.empty((n_cells, len(keys)))
    got = _I64(-1)
    if n_nodes < 0:
        _check(lib().titan_checkpoint_read(os.fsencode(path), n_cells, len(keys), kp, _ptr(vals), -1, None,
                                           ctypes.byref(got)), "titan_checkpoint_read")
        n_nodes = got.value
    xyz = np.empty((max(n_nodes, 0), 3))
    _check(lib().titan_checkpoint_read(os.fsencode(path), n_cells, len(keys), kp, _ptr(vals), n_nodes,
                                       _ptr(xyz) if n_nodes > 0 else None, ctypes.byref(got)),
           "titan_checkpoint_read")
    return vals, (xyz if got.value > 0 else None)


def mk_s_finish_batch(vol, sm, sp, w_r, a2, kn):
    """Cell::MK_normir_Moments + Cell::MK_calc_Sm on the device for raw tallies sm [n, n_w],
    sp [n, n_w, 2] and cell volumes vol [n]: (sm_out, sp_out)."""
    sm = _c(sm, np.float64)
    sp = _c(sp, np.float64)
    n, nw = sm.shape
    if sp.shape != (n, nw, 2):
        raise ValueError("sp must be [n, n_w, 2]")
    so, po = np.zeros_like(sm), np.zeros_like(sp)
    _check(lib().titan_mk_s_finish_batch(n, nw, float(w_r), float(a2), float(kn), _ptr(_c(vol, np.float64)), _ptr(sm),
                                         _ptr(sp), _ptr(so), _ptr(po)), "titan_mk_s_finish_batch")
    return so, po


def mk_s_write(path, sm, sp):
    """Cell::write_S_ToFile for one cell: sm [n_w], sp [n_w, 2]."""
    sm = _c(sm, np.float64).reshape(-1)
    sp = _c(sp, np.float64).reshape(sm.shape[0], 2)
    _check(lib().titan_mk_s_write(os.fsencode(path), sm.shape[0], _ptr(sm), _ptr(sp)), "titan_mk_s_write")


def mk_s_read(path, n_w):
    """Cell::read_S_FromFile for one cell: (sm [n_w], sp [n_w, 2])."""
    sm, sp = np.zeros(int(n_w)), np.zeros((int(n_w), 2))
    _check(lib().titan_mk_s_read(os.fsencode(path), int(n_w), _ptr(sm), _ptr(sp)), "titan_mk_s_read")
    return sm, sp


MK_PARAMETERS = ("MK_n_H", "MK_n_H1", "MK_n_H2", "MK_n_H3", "MK_n_H4")


def mk_parameters_write(path, values, keys=MK_PARAMETERS):
    """Setka::Save_cell_MK_parameters file: values [n_cells, len(keys)] (titan_mk_go's density by
    default: MK_n_H, MK_n_H1..4)."""
    vals = _c(values, np.float64).reshape(-1, len(keys))
    arr, kp = _keys(keys)
    _check(lib().titan_mk_parameters_write(os.fsencode(path), vals.shape[0], len(keys), kp, _ptr(vals)),
           "titan_mk_parameters_write")


def mk_parameters_read(path, values, keys=MK_PARAMETERS, cell_mk=None, zone_except=-1):
    """Setka::Download_cell_MK_parameters: overwrites values [n_cells, len(keys)] (float64,
    C-contiguous, in place) with the file's, except in the cells of MK zone zone_except (cell_mk
    from Set_MK_Zone); returns how many of the keys the file holds."""
    if not (isinstance(values, np.ndarray) and values.dtype == np.float64 and values.flags.c_contiguous
            and values.ndim == 2 and values.shape[1] == len(keys)):
        raise ValueError("values must be a C-contiguous float64 array [n_cells, len(keys)]")
    arr, kp = _keys(keys)
    cm = None if cell_mk is None else _c(cell_mk, np.int8)
    if cm is not None and cm.shape[0] != values.shape[0]:
        raise ValueError("cell_mk must hold one zone per cell")
    got = _I32(0)
    _check(lib().titan_mk_parameters_read(os.fsencode(path), values.shape[0], len(keys), kp, _ptr(values),
                                          _ptr(cm), zone_except, ctypes.byref(got)), "titan_mk_parameters_read")
    return got.value


# ---- mesh helpers ---------------------------------------------------------------------------
def mesh_connect(cell_nodes, n_nodes):
    """Setka::New_connect (Setka.cpp:2066-2249): returns face_nodes, face_cells, cell_faces."""
    cn = _c(cell_nodes, np.int32).reshape(-1)
    nc = cn.size // 8
    fnodes = np.empty(6 * nc * 4, np.int32)
    fcells = np.empty(6 * nc * 2, np.int32)
    cfaces = np.empty(nc * 6, np.int32)
    nf = lib().titan_mesh_connect(nc, _ptr(cn), int(n_nodes), _ptr(fnodes), _ptr(fcells), _ptr(cfaces))
    if nf < 0:
        raise TitanError(int(nf), "titan_mesh_connect")
    return fnodes[:4 * nf].reshape(nf, 4).copy(), fcells[:2 * nf].reshape(nf, 2).copy(), cfaces.reshape(nc, 6)


def mesh_cube(n, jitter, seed):
    xyz = np.empty(((n + 1) ** 3, 3))
    cn = np.empty((n ** 3, 8), np.int32)
    _check(lib().titan_mesh_cube(n, jitter, seed, _ptr(xyz), _ptr(cn)), "titan_mesh_cube")
    return xyz, cn


def mesh_shell(n_edge, radii):
    radii = _c(radii, np.float64)
    nl = radii.size - 1
    ns = 6 * n_edge * n_edge + 2
    xyz = np.empty(((nl + 1) * ns, 3))
    cn = np.empty((nl * 6 * n_edge * n_edge, 8), np.int32)
    _check(lib().titan_mesh_shell(n_edge, nl, _ptr(radii), _ptr(xyz), _ptr(cn)), "titan_mesh_shell")
    return xyz, cn


# ---- Riemann solver -----------------------------------------------------------------------------
def chlld(n_state, normal, qL, qR, w=None, n_disc=None, flags=None, kL=None, kR=None, precision=PREC_FP64):
    """Phys_param::chlld over a batch.  Returns (q [n,8], k [n,K], ds [n,3], branch [n], rc)."""
    qL = _c(qL, np.float64).reshape(-1, 8)
    qR = _c(qR, np.float64).reshape(-1, 8)
    n = qL.shape[0]
    ns = _c(np.broadcast_to(n_state, (n,)), np.int8)
    nd = None if n_disc is None else _c(np.broadcast_to(n_disc, (n,)), np.int8)
    fl = None if flags is None else _c(np.broadcast_to(flags, (n,)), np.uint8)
    nrm = _c(normal, np.float64).reshape(n, 3)
    ww = None if w is None else _c(np.broadcast_to(w, (n,)), np.float64)
    K = 0 if kL is None else int(np.asarray(kL).reshape(n, -1).shape[1])
    kLa = None if K == 0 else _c(kL, np.float64).reshape(n, K)
    kRa = None if K == 0 else _c(kR, np.float64).reshape(n, K)
    q = np.zeros((n, 8))
    k = np.zeros((n, max(K, 1)))
    ds = np.zeros((n, 3))
    br = np.zeros(n, np.uint8)
    rc = lib().titan_chlld_batch(n, _ptr(ns), _ptr(nd), _ptr(fl), _ptr(nrm), _ptr(ww), _ptr(qL), _ptr(qR), K,
                                 _ptr(kLa), _ptr(kRa), _ptr(q), _ptr(k) if K else None, _ptr(ds), _ptr(br),
                                 int(precision))
    if rc not in (TITAN_OK, TITAN_E_RIEMANN):
        _check(rc, "titan_chlld_batch")
    return q, k[:, :K], ds, br, rc


def cx_exact_batch(in10, nodes=100):
    a = _c(in10, np.float64).reshape(-1, 10)
    out = np.zeros((a.shape[0], 5))
    _check(lib().titan_cx_exact_batch(a.shape[0], _ptr(a), int(nodes), _ptr(out)), "titan_cx_exact_batch")
    return out


# ---- context ---------------------------------------------------------------------------------
def sphere_fit(points):
    """Smooth_head_TS3's sphere centre of n >= 3 points [n, 3] (host routine of the library)."""
    P = np.ascontiguousarray(points, np.float64).reshape(-1, 3)
    c = np.zeros(3)
    _check(lib().titan_sphere_fit(P.shape[0], _ptr(P), _ptr(c)), "titan_sphere_fit")
    return c


def circle_fit(points):
    """fitCircleRobust's (centre [2], radius) of n >= 3 points [n, 2] (host routine of the library)."""
    P = np.ascontiguousarray(points, np.float64).reshape(-1, 2)
    c = np.zeros(3)
    _check(lib().titan_circle_fit(P.shape[0], _ptr(P), _ptr(c)), "titan_circle_fit")
    return c[:2], float(c[2])


class Setka:
    """One device context: mesh, both time layers of state and geometry, step scalars."""

    def __init__(self, mesh, params):
        self._keep = []
        m = MeshC()
        m.n_nodes, m.n_cells, m.n_faces = mesh.n_nodes, mesh.n_cells, mesh.n_faces

        def keep(a, dt):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=dt).reshape(-1)
            self._keep.append(a)
            return a.ctypes.data

        m.cell_nodes = keep(mesh.cell_nodes, np.int32)
        m.face_nodes = keep(mesh.face_nodes, np.int32)
        m.face_cells = keep(mesh.face_cells, np.int32)
        m.cell_faces = keep(mesh.cell_faces, np.int32)
        m.face_bc = keep(mesh.face_bc, np.int8)
        m.face_method = keep(mesh.face_method, np.int8)
        m.face_flags = keep(mesh.face_flags, np.uint8)
        m.face_bstate = keep(mesh.face_bstate, np.float64)
        m.cell_zone = keep(mesh.cell_zone, np.int8)
        m.cell_is_tvd = keep(mesh.cell_is_tvd, np.uint8)
        m.node_xyz = keep(mesh.node_xyz, np.float64)
        m.node_disp = keep(mesh.node_disp, np.float64)
        m.face_bscalar = keep(getattr(mesh, "face_bscalar", None), np.float64)
        p = ParamsC()
        for k, _ in ParamsC._fields_:
            setattr(p, k, params.get(k, _PARAM_DEFAULTS[k]) if k in _PARAM_DEFAULTS else params[k])
        self.n_cells, self.n_faces, self.n_nodes = mesh.n_cells, mesh.n_faces, mesh.n_nodes
        self.params = dict(params)
        ctx = ctypes.c_void_p()
        _check(lib().titan_create(ctypes.byref(ctx), ctypes.byref(m), ctypes.byref(p)), "titan_create")
        self._keep = []
        self.ctx = ctx

    def close(self):
        if getattr(self, "ctx", None):
            lib().titan_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_state(self, prim, layer=0):
        a = _c(prim, np.float64).reshape(self.n_cells, 8)
        _check(lib().titan_set_state(self.ctx, _ptr(a), layer), "titan_set_state", self.ctx)

    def get_state(self, layer=None):
        if layer is None:
            layer = self.current_layer()
        a = np.empty((self.n_cells, 8))
        _check(lib().titan_get_state(self.ctx, _ptr(a), layer), "titan_get_state", self.ctx)
        return a

    def set_scalars(self, qhe, layer=0):
        """Passive scalars [n_cells, 2] = (Q, rho_He) of a layer (params passive_scalars = 1)."""
        a = _c(qhe, np.float64).reshape(self.n_cells, 2)
        _check(lib().titan_set_scalars(self.ctx, _ptr(a), layer), "titan_set_scalars", self.ctx)

    def get_scalars(self, layer=None):
        if layer is None:
            layer = self.current_layer()
        a = np.empty((self.n_cells, 2))
        _check(lib().titan_get_scalars(self.ctx, _ptr(a), layer), "titan_get_scalars", self.ctx)
        return a

    def set_nodes(self, xyz, layer):
        a = _c(xyz, np.float64).reshape(self.n_nodes, 3)
        _check(lib().titan_set_nodes(self.ctx, _ptr(a), layer), "titan_set_nodes", self.ctx)

    def get_nodes(self, layer):
        a = np.empty((self.n_nodes, 3))
        _check(lib().titan_get_nodes(self.ctx, _ptr(a), layer), "titan_get_nodes", self.ctx)
        return a

    def set_neutrals(self, nh5):
        a = _c(nh5, np.float64).reshape(self.n_cells, 5)
        _check(lib().titan_set_neutrals(self.ctx, _ptr(a)), "titan_set_neutrals", self.ctx)

    def set_h_fluids(self, h):
        """Hydrogen fluids [n_cells, n_h, 5] = (rho_Hk, p_Hk, V_Hk) of the closed-form source."""
        a = _c(h, np.float64).reshape(self.n_cells, -1, 5)
        _check(lib().titan_set_h_fluids(self.ctx, a.shape[1], _ptr(a)), "titan_set_h_fluids", self.ctx)

    def Calculating_measure(self, layer):
        _check(lib().titan_calculating_measure(self.ctx, layer), "titan_calculating_measure", self.ctx)

    def geometry(self, layer):
        nc, nf = self.n_cells, self.n_faces
        cc, cv = np.empty((nc, 3)), np.empty(nc)
        fc, fn, fa = np.empty((nf, 3)), np.empty((nf, 3)), np.empty(nf)
        _check(lib().titan_get_geometry(self.ctx, layer, _ptr(cc), _ptr(cv), _ptr(fc), _ptr(fn), _ptr(fa)),
               "titan_get_geometry", self.ctx)
        return dict(cell_center=cc, cell_volume=cv, face_center=fc, face_normal=fn, face_area=fa)

    def Culc_Gran_Potok(self, now, dt, want_branch=True, want_scalars=False):
        """The face loop: returns (flux [nf,9], dt_min, branch [nf]), and with want_scalars the
        passive scalar fluxes [nf,2] (PQ, Prho_He) as a fourth item."""
        flux = np.empty((self.n_faces, 9))
        kflux = np.empty((self.n_faces, 2)) if want_scalars else None
        dtm = ctypes.c_double(0.0)
        br = np.empty(self.n_faces, np.uint8) if want_branch else None
        _check(lib().titan_gran_potok(self.ctx, now, float(dt), _ptr(flux), _ptr(kflux), ctypes.byref(dtm), _ptr(br)),
               "titan_gran_potok", self.ctx)
        if want_scalars:
            return flux, dtm.value, br, kflux
        return flux, dtm.value, br

    def Go(self, steps, is_inner_area=None):
        """titan_go: `steps` steps of every cell, the dt schedule continuing across calls; with
        is_inner_area (True / False), titan_go_area: Setka::Go(is_inner_area, steps) on one area,
        the dt schedule restarting as in the reference."""
        st = StatsC()
        if is_inner_area is None:
            _check(lib().titan_go(self.ctx, int(steps), ctypes.byref(st)), "titan_go", self.ctx)
        else:
            _check(lib().titan_go_area(self.ctx, 1 if is_inner_area else 0, int(steps), ctypes.byref(st)),
                   "titan_go_area", self.ctx)
        return {k: getattr(st, k) for k, _ in StatsC._fields_}

    def set_inner_area(self, cell_is_inner):
        """Cell::is_inner per cell (Setka::Init_boundary_grans)."""
        a = np.ascontiguousarray(cell_is_inner, np.uint8).reshape(-1)
        _check(lib().titan_set_inner_area(self.ctx, _ptr(a)), "titan_set_inner_area", self.ctx)

    def current_layer(self):
        return lib().titan_current_layer(self.ctx)

    def Save_cell_parameters(self, path, extra=None):
        """Setka::Save_cell_parameters: the 8 plasma keys of the newest layer, extra keys
        {name: [n_cells]} from the host, and the newest nodes."""
        extra = extra or {}
        names = list(extra)
        vals = np.ascontiguousarray(np.stack([np.asarray(extra[k], np.float64) for k in names], axis=1)
                                    if names else np.zeros((self.n_cells, 0)))
        arr, kp = _keys(names) if names else (None, None)
        _check(lib().titan_save_cell_parameters(self.ctx, os.fsencode(path), len(names), kp,
                                                _ptr(vals) if names else None),
               "titan_save_cell_parameters", self.ctx)

    def Download_cell_parameters(self, path):
        _check(lib().titan_download_cell_parameters(self.ctx, os.fsencode(path)), "titan_download_cell_parameters",
               self.ctx)

    def Init_TVD(self):
        _check(lib().titan_init_tvd(self.ctx), "titan_init_tvd", self.ctx)

    # ---- fitted surfaces (Setka_move.cpp, Luch.cpp) ----
    def set_rays(self, rays):
        """rays: dict with ray_nodes [n_rays, nodes_per_ray] (reference node ids, inner to outer),
        ray_angles [n_rays, 2] (the, phi), R0, R1, R5, M0, M1, M11, M2, M3, M4, da1..da5, phi_a."""
        nodes = np.ascontiguousarray(rays["ray_nodes"], np.int32)
        ang = np.ascontiguousarray(rays["ray_angles"], np.float64)
        r = RaysC()
        r.n_rays, r.nodes_per_ray = nodes.shape[0], nodes.shape[1]
        r.ray_nodes, r.ray_angles = _ptr(nodes), _ptr(ang)
        for k in ("R0", "R1", "R5", "M0", "M1", "M11", "M2", "M3", "M4", "da1", "da2", "da3", "da4", "da5", "phi_a"):
            setattr(r, k, rays[k])
        _check(lib().titan_set_rays(self.ctx, ctypes.byref(r)), "titan_set_rays", self.ctx)

    def set_rays_ext(self, types, node_lists, opor, angles, geo):
        """The B, C, D, E, H and G rays: types [n] (RAY_B ... RAY_G), node_lists (reference node ids
        of each ray, inner to outer), opor [n, 4] support nodes (-1 unused), angles [n, 2] (the, phi),
        geo: L6, L7, N4, N5, MF, dd1..dd8, ba1..ba5, md1..md3, ea1..ea3 (Geo_param)."""
        t = np.ascontiguousarray(types, np.int8)
        off = np.zeros(len(node_lists) + 1, np.int64)
        off[1:] = np.cumsum([len(x) for x in node_lists])
        nodes = np.ascontiguousarray(np.concatenate([np.asarray(x, np.int32) for x in node_lists])
                                     if len(node_lists) else np.zeros(1, np.int32), np.int32)
        op = np.ascontiguousarray(opor, np.int32).reshape(-1, 4)
        ang = np.ascontiguousarray(angles, np.float64).reshape(-1, 2)
        r = RaysExtC()
        r.n_rays = len(t)
        r.ray_type, r.ray_off, r.ray_nodes, r.ray_opor, r.ray_angles = (_ptr(t), _ptr(off), _ptr(nodes), _ptr(op),
                                                                        _ptr(ang))
        for k, _ in RaysExtC._fields_[6:]:
            setattr(r, k, geo[k])
        self._rays_ext_keep = (t, off, nodes, op, ang)
        _check(lib().titan_set_rays_ext(self.ctx, ctypes.byref(r)), "titan_set_rays_ext", self.ctx)

    def Smooth_angle_HP(self, nodes, nb6):
        """Setka::Smooth_angle_HP: circle centres from node layer 0 and the azimuthal neighbours nb6
        [n, 6] (i-1, i-2, i+1, i+2, i+3, i-3) of HP nodes."""
        nd = np.ascontiguousarray(nodes, np.int32)
        nb = np.ascontiguousarray(nb6, np.int32).reshape(-1, 6)
        _check(lib().titan_smooth_angle_hp(self.ctx, len(nd), _ptr(nd), _ptr(nb)), "titan_smooth_angle_hp",
               self.ctx)

    def set_hp_grid(self, nodes=(), nb8=None, tail=None, N4=0, bs_ref=(), bs_lists=(), **coef):
        """HP grid smoothing (nodes, nb8 [n, 8]: AA1, AA11, AA3, AA33, AA2, AA22, AA4, AA44), the tail
        correction (tail [n_az, len] D-ray HP nodes, N4) and the BS correction (bs_ref [n_az],
        bs_lists: the BS nodes of each azimuth); coef: sglag_HP_k_angle, _k, _angle, _along."""
        k = dict(HP_GRID_DEFAULTS)
        k.update(coef)
        nd = np.ascontiguousarray(nodes, np.int32)
        nb = np.ascontiguousarray(nb8 if nb8 is not None else np.zeros((0, 8)), np.int32).reshape(-1, 8)
        tl = np.ascontiguousarray(tail if tail is not None else np.zeros((0, 0)), np.int32)
        br = np.ascontiguousarray(bs_ref, np.int32)
        off = np.zeros(len(bs_lists) + 1, np.int32)
        off[1:] = np.cumsum([len(x) for x in bs_lists])
        bn = np.ascontiguousarray(np.concatenate([np.asarray(x, np.int32) for x in bs_lists])
                                  if len(bs_lists) else np.zeros(1), np.int32)
        g = HpGridC()
        g.n_grid, g.grid_node, g.grid_nb = len(nd), _ptr(nd), _ptr(nb)
        for key in HP_GRID_DEFAULTS:
            setattr(g, key, k[key])
        g.n_tail = tl.shape[0] if tl.ndim == 2 else 0
        g.tail_len = tl.shape[1] if tl.ndim == 2 else 0
        g.N4, g.tail = int(N4), _ptr(tl)
        g.n_bs, g.bs_ref, g.bs_off, g.bs_nodes = len(br), _ptr(br), _ptr(off), _ptr(bn)
        self._hp_keep = (nd, nb, tl, br, off, bn)
        _check(lib().titan_set_hp_grid(self.ctx, ctypes.byref(g)), "titan_set_hp_grid", self.ctx)

    def Smooth_head_TS3(self):
        _check(lib().titan_smooth_head_ts3(self.ctx), "titan_smooth_head_ts3", self.ctx)

    def Smooth_head_HP3(self):
        _check(lib().titan_smooth_head_hp3(self.ctx), "titan_smooth_head_hp3", self.ctx)

    def Find_Yzel_Sosed_for_BS(self):
        _check(lib().titan_find_yzel_sosed_for_bs(self.ctx), "titan_find_yzel_sosed_for_bs", self.ctx)

    def Culc_Velocity_surface(self, now, time):
        """Moves the TS/HP/BS nodes of layer `now` into layer 1-now; returns the averaged velocity
        of every node [n_nodes, 3] (0 off the surfaces)."""
        v = np.zeros((self.n_nodes, 3))
        _check(lib().titan_culc_velocity_surface(self.ctx, int(now), float(time), _ptr(v)),
               "titan_culc_velocity_surface", self.ctx)
        return v

    def dvigenie(self, layer):
        """Luch::dvigenie(layer) of every ray."""
        _check(lib().titan_dvigenie(self.ctx, int(layer)), "titan_dvigenie", self.ctx)

    def tvd_cells(self):
        """Setka::Init_TVD neighbours [n_faces, 2] (AA behind cells[0], BB behind cells[1], -1 = none)."""
        out = np.empty((self.n_faces, 2), np.int32)
        _check(lib().titan_tvd_cells(self.ctx, _ptr(out)), "titan_tvd_cells", self.ctx)
        return out

    def sources(self, layer, mode):
        out = np.empty((self.n_cells, 5))
        _check(lib().titan_sources(self.ctx, layer, mode, _ptr(out)), "titan_sources", self.ctx)
        return out

    def sources_h(self, layer, n_h, arise=None):
        """Calc_sourse_MF_Bera with the hydrogen fluids' outputs: (plasma [n_cells, 5], fluids
        [n_cells, n_h, 5]); arise [4, n_h] 1-based target fluid per zone, None for fluid = zone."""
        out = np.empty((self.n_cells, 5))
        hout = np.empty((self.n_cells, n_h, 5))
        ar = None if arise is None else np.ascontiguousarray(arise, np.int8).reshape(4, n_h)
        _check(lib().titan_sources_h(self.ctx, layer, _ptr(ar), _ptr(out), _ptr(hout)), "titan_sources_h", self.ctx)
        return out, hout

    def MK_go(self, n_per_face, seed=1, max_steps=100000, layer=None):
        """Setka::MK_go (Setka_MK.cpp:1225-2316) for the interstellar hydrogen: returns (density
        [n_cells, 5] = MK_n_H, MK_n_H1..MK_n_H4 per unit volume, stats dict)."""
        layer = self.current_layer() if layer is None else layer
        out = np.empty((self.n_cells, 5))
        st = np.zeros(8, np.int64)
        _check(lib().titan_mk_go(self.ctx, int(layer), int(n_per_face), int(seed) & (2**64 - 1), int(max_steps),
                                 _ptr(out), _ptr(st)), "titan_mk_go", self.ctx)
        return out, dict(zip(MK_STATS, (int(v) for v in st)))

    def Set_MK_Zone(self, r=1):
        """Setka::Set_MK_Zone (Setka_MK.cpp:142-683) and the zone borders' distributions with r-fold
        refined bins: returns (MK zone per cell [n_cells], zone-list mask per face [n_faces])."""
        cm = np.zeros(self.n_cells, np.int8)
        fm = np.zeros(self.n_faces, np.uint8)
        _check(lib().titan_set_mk_zone(self.ctx, int(r), _ptr(cm), _ptr(fm)), "titan_set_mk_zone", self.ctx)
        self._mk_r = int(r)
        return cm, fm

    def MK_go_zone(self, zone, n_per_face, seed=1, max_steps=100000, layer=None, density=None):
        """MK_prepare(zone) + MK_go(zone): the zone's cells of `density` ([n_cells, 5], created as
        zeros when None) get their densities; returns (density, stats dict)."""
        layer = self.current_layer() if layer is None else layer
        out = np.zeros((self.n_cells, 5)) if density is None else density
        assert out.shape == (self.n_cells, 5) and out.dtype == np.float64 and out.flags.c_contiguous
        st = np.zeros(9, np.int64)
        _check(lib().titan_mk_go_zone(self.ctx, int(layer), int(zone), int(n_per_face), int(seed) & (2**64 - 1),
                                      int(max_steps), _ptr(out), _ptr(st)), "titan_mk_go_zone", self.ctx)
        return out, dict(zip(MK_ZONE_STATS, (int(v) for v in st)))

    def mk_dist(self, face, sort, side, w=None):
        """A listed interior face's distribution (108 r^3 bins): read it, or with w write it."""
        n = 108 * self._mk_r ** 3
        if w is None:
            out = np.zeros(n)
            _check(lib().titan_mk_dist_get(self.ctx, int(face), int(sort), int(side), _ptr(out)), "titan_mk_dist_get",
                   self.ctx)
            return out
        a = np.ascontiguousarray(w, np.float64).reshape(n)
        _check(lib().titan_mk_dist_set(self.ctx, int(face), int(sort), int(side), _ptr(a)), "titan_mk_dist_set",
               self.ctx)
        return a

    def MK_source_S(self, n_w, w_r, arise=None):
        """Turns on (n_w > 0) the pickup-ion source tallies S-, S+ of MK_source_S for the next
        MK_go / MK_go_zone calls: n_w speed bins on (0, w_r) (pui_nW, pui_wR); arise [4, 4] (zone,
        sort) -> pickup-ion fluid 0..2, None for the reference's tables with pickup ions."""
        a = None if arise is None else np.ascontiguousarray(arise, np.int8).reshape(16)
        _check(lib().titan_set_mk_source_s(self.ctx, int(n_w), float(w_r), _ptr(a)), "titan_set_mk_source_s",
               self.ctx)
        self._pui_nw = int(n_w)

    def mk_S(self):
        """The last Monte-Carlo call's (S- [n_cells, n_w], S+ [n_cells, n_w, 2]) after
        MK_normir_Moments and MK_calc_Sm."""
        nw = getattr(self, "_pui_nw", 0)
        sm = np.zeros((self.n_cells, max(nw, 1)))
        sp = np.zeros((self.n_cells, max(nw, 1), 2))
        _check(lib().titan_mk_get_s(self.ctx, _ptr(sm), _ptr(sp)), "titan_mk_get_s", self.ctx)
        return sm, sp

    def set_timing(self, on=True):
        _check(lib().titan_set_timing(self.ctx, 1 if on else 0), "titan_set_timing", self.ctx)

    def kernel_ms(self, which):
        return lib().titan_kernel_ms(self.ctx, which)

    def tile_info(self):
        a = np.zeros(8, np.int64)
        _check(lib().titan_tile_info(self.ctx, _ptr(a)), "titan_tile_info", self.ctx)
        keys = ("tile_cells", "max_nodes", "max_faces", "lds_bytes", "lds_limit", "tiles", "face_records",
                "tvd_faces")
        return {k: int(v) for k, v in zip(keys, a)}

    # ---- surrogate ----
    def mlp_set(self, params, in_norm16, out_scale5, box16):
        P = _c(params, np.float32)
        a, b, c = _c(in_norm16, np.float32), _c(out_scale5, np.float32), _c(box16, np.float32)
        _check(lib().titan_mlp_set(self.ctx, _ptr(P), _ptr(a), _ptr(b), _ptr(c)), "titan_mlp_set", self.ctx)

    def mlp_get(self):
        P = np.empty(lib().titan_mlp_n_params(), np.float32)
        _check(lib().titan_mlp_get(self.ctx, _ptr(P)), "titan_mlp_get", self.ctx)
        return P

    def mlp_forward(self, x):
        x = _c(x, np.float32).reshape(-1, 8)
        y = np.empty((x.shape[0], 5), np.float32)
        _check(lib().titan_mlp_forward(self.ctx, x.shape[0], _ptr(x), _ptr(y)), "titan_mlp_forward", self.ctx)
        return y

    def mlp_loss_grad(self, x, y):
        x = _c(x, np.float32).reshape(-1, 8)
        y = _c(y, np.float32).reshape(-1, 5)
        loss = ctypes.c_float(0)
        g = np.empty(lib().titan_mlp_n_params(), np.float32)
        _check(lib().titan_mlp_loss_grad(self.ctx, x.shape[0], _ptr(x), _ptr(y), ctypes.byref(loss), _ptr(g)),
               "titan_mlp_loss_grad", self.ctx)
        return loss.value, g

    def mlp_train(self, x, y, epochs, batch, lr, seed=0):
        x = _c(x, np.float32).reshape(-1, 8)
        y = _c(y, np.float32).reshape(-1, 5)
        hist = np.zeros(max(1, epochs), np.float32)
        _check(lib().titan_mlp_train(self.ctx, x.shape[0], _ptr(x), _ptr(y), int(epochs), int(batch), float(lr),
                                     int(seed), _ptr(hist)), "titan_mlp_train", self.ctx)
        return hist[:epochs]
