"""ctypes loader of the CPU oracle (liboracle.so).  TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg import this module, and only
as the checker or the timed CPU baseline; the product path never does.
"""
import ctypes
import math
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")

_P = ctypes.c_void_p
_D = ctypes.c_double
_I = ctypes.c_int
_I64 = ctypes.c_int64


class OrMesh(ctypes.Structure):
    _fields_ = [("n_nodes", _I64), ("n_cells", _I64), ("n_faces", _I64), ("cell_nodes", _P), ("cell_faces", _P),
                ("face_nodes", _P), ("face_cells", _P), ("face_bc", _P), ("face_method", _P), ("face_flags", _P),
                ("face_bstate", _P), ("cell_zone", _P), ("cell_is_tvd", _P), ("face_tvd", _P), ("cell_inner", _P),
                ("face_bscal", _P)]


class OrParams(ctypes.Structure):
    _fields_ = [("gamma", _D), ("kfl", _D), ("v_inf", _D), ("move_mesh", ctypes.c_int32), ("n_disc", ctypes.c_int32),
                ("motion_omega", _D), ("motion_phase", _D), ("tvd", ctypes.c_int32), ("snos_on_hp", ctypes.c_int32),
                ("par_a_2", _D), ("par_n_H_LISM", _D), ("par_Kn", _D), ("area", ctypes.c_int32)]

# parameters a params dict may lack (Phys_param.h:99-101 defaults for par_*; area 0: every cell)
PARAM_DEFAULTS = {"tvd": 0, "snos_on_hp": 0, "par_a_2": 0.102046, "par_n_H_LISM": 3.0, "par_Kn": 40.0906,
                  "area": 0}


class OrGeom(ctypes.Structure):
    _fields_ = [("cell_center", _P), ("cell_volume", _P), ("face_center", _P), ("face_normal", _P),
                ("face_area", _P)]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        sig = {
            "or_chlld": (_I, [_I, _I, _I, _D, _D, _D, _D, _P, _P, _I, _P, _P, _P, _P, _P, _P, _P, _P]),
            "or_chlld_batch": (_I, [_I64, _P, _P, _P, _P, _P, _P, _P, _I, _P, _P, _P, _P, _P, _P]),
            "or_get_bazis": (None, [_P, _P, _P]),
            "or_get_potok": (None, [_D] * 12 + [_P]),
            "or_measure": (None, [_P, _P, _P]),
            "or_measure_gcl": (None, [_P, _P, _P]),
            "or_det_cos": (_D, [_D]),
            "or_face_velocities": (None, [_P, _P, _P, _P, _P, _D, _I, _P]),
            "or_motion": (None, [_I64, _P, _P, _D, _D, _D, _D, _P]),
            "or_face_flux": (_I, [_P, _P, _P, _P, _P, _P, _D, _P, _P, _P]),
            "or_face_states": (_I, [_P, _P, _I64, _P, _P, _P, _P]),
            "or_init_tvd": (None, [_P, _P, _P]),
            "or_minmod": (_D, [_D, _D]),
            "or_linear": (_D, [_D] * 7),
            "or_update": (_I, [_P, _P, _P, _P, _P, _P, _P, _D, _P, _P, _P]),
            "or_go": (_I, [_P, _P, _P, _P, _P, _I64, _I, _I, _P, _P, _P, _I, _P, _P]),
            "or_culc_velocity_surface": (_I, [_P, _P, _P, _P, _P, _P, _D, _P, _P]),
            "or_dvigenie": (None, [_P, _P]),
            "or_dvigenie_ext": (None, [_P, _P]),
            "or_circle_fit": (_D, [_I64, _P, _P]),
            "or_hp_terms": (_I, [_P, _P, _D, _D, _P, ctypes.c_int32, _D, _D, _D, _D, _P]),
            "or_mk_setup": (_I64, [_P, _P, _D, ctypes.c_int32, _P, _P, _P]),
            "or_mk_go": (_I, [_P, _P, _P, _P, _P, _P, _P]),
            "or_mk_change_velosity_batch": (None, [_I64, ctypes.c_uint64, _D, _P, _P, _D, _I, _P, _P]),
            "or_set_mk_zone": (_I, [_P, _P, _I, _P, _P]),
            "or_plasma_components_1": (None, [_D, _D, _D, _I, _P]),
            "or_mk_pui_finish": (None, [ctypes.c_int64, ctypes.c_int32, _D, _D, _D, _P, _P, _P, _P, _P]),
            "or_mk_setup_zone": (_I64, [_P, _P, _D, ctypes.c_int32, ctypes.c_int32, _P, _P, _P, _P, _P, _P, _P, _P]),
            "or_mk_deposit": (None, [_P, ctypes.c_int32, _I, _I, _P, _P, _D]),
            "or_mk_sample_batch": (None, [_P, ctypes.c_int32, _I, _I, _P, _I64, ctypes.c_uint64, _P]),
            "or_bera_source": (None, [_P, _I, _P, _D, _D, _D, _P]),
            "or_new_connect": (_I64, [_I64, _P, _I64, _P, _P, _P]),
            "or_cx_rate_ref": (_D, [_D, _D, _D, _I, _I]),
            "or_sigma_ls": (_D, [_D]),
            "or_cx_source": (None, [_D, _D, _P, _D, _D, _P, _I, _I, _P]),
            "or_mlp_n_params": (_I64, []),
            "or_mlp_forward": (None, [_P, _I64, _P, _P]),
            "or_mlp_forward_bf16": (None, [_P, _I64, _P, _P]),
            "or_mlp_loss_grad": (_D, [_P, _I64, _P, _P, _P]),
            "or_set_mistake": (None, [_I]),
            "or_face_states_k": (_I, [_P, _P, _I64, _P, _P, _P, _P, _P, _P, _P]),
            "or_face_flux_k": (_I, [_P, _P, _P, _P, _P, _P, _P, _D, _P, _P, _P, _P]),
            "or_update_k": (_I, [_P, _P, _P, _P, _P, _P, _P, _P, _P, _D, _P, _P, _P, _P]),
            "or_go_k": (_I, [_P, _P, _P, _P, _P, _P, _I64, _I, _I, _P, _P, _P, _I, _P, _P]),
            "or_determ_zone": (_I, [_D, _D, _I]),
            "or_bera_source_k": (None, [_P, _D, _D, _I, _I, _P, _D, _D, _D, _P]),
            "or_bera_source_h": (None, [_P, _D, _D, _I, _I, _P, _D, _D, _D, _P, _P, _P]),
        }
        for k, (r, a) in sig.items():
            f = getattr(L, k)
            f.restype = r
            f.argtypes = a
        _lib = L
    return _lib


# planted mistakes of the restatement (OR_MISTAKE_* of titan_oracle.h), for tests that show a pin
# catches a plausible error
MISTAKE_POWELL_MOM_8PI = 1 << 0
MISTAKE_POWELL_IND_SIGN = 1 << 1
MISTAKE_POWELL_ENERGY = 1 << 2
MISTAKE_PDIVB_HALF = 1 << 3
MISTAKE_ZONE2_8PI = 1 << 4
MISTAKE_HLLD_VZZ_SIGN = 1 << 5
MISTAKE_HLLD_BZZ_WEIGHTS = 1 << 6
MISTAKE_HLLD_TB_BN2 = 1 << 7
MISTAKE_SCAL_ZONE1_PLAIN = 1 << 8
MISTAKE_SCAL_TVD_R2_KEPT = 1 << 9
MISTAKE_HE_CLIP_RHO = 1 << 10


class planted_mistake:
    """Context manager: the oracle runs with the given OR_MISTAKE_* bits, then with none."""

    def __init__(self, bits):
        self.bits = int(bits)

    def __enter__(self):
        lib().or_set_mistake(self.bits)
        return self

    def __exit__(self, *exc):
        lib().or_set_mistake(0)
        return False


def _p(a):
    return None if a is None else a.ctypes.data


class OracleMesh:
    """Holds contiguous copies of a configs.Mesh and its ctypes view.  tvd=True also computes the
    TVD neighbours of every face from the initial geometry (Setka::Init_TVD)."""

    def __init__(self, mesh, tvd=False):
        self.m = mesh
        self.arrs = dict(
            cell_nodes=np.ascontiguousarray(mesh.cell_nodes, np.int32), cell_faces=np.ascontiguousarray(mesh.cell_faces, np.int32),
            face_nodes=np.ascontiguousarray(mesh.face_nodes, np.int32), face_cells=np.ascontiguousarray(mesh.face_cells, np.int32),
            face_bc=np.ascontiguousarray(mesh.face_bc, np.int8), face_method=np.ascontiguousarray(mesh.face_method, np.int8),
            face_flags=None if mesh.face_flags is None else np.ascontiguousarray(mesh.face_flags, np.uint8),
            face_bstate=None if mesh.face_bstate is None else np.ascontiguousarray(mesh.face_bstate, np.float64),
            cell_zone=np.ascontiguousarray(mesh.cell_zone, np.int8), cell_is_tvd=np.ascontiguousarray(mesh.cell_is_tvd, np.uint8),
            face_tvd=None, cell_inner=None,
            face_bscal=None if getattr(mesh, "face_bscalar", None) is None
            else np.ascontiguousarray(mesh.face_bscalar, np.float64))
        c = OrMesh()
        c.n_nodes, c.n_cells, c.n_faces = mesh.n_nodes, mesh.n_cells, mesh.n_faces
        for k, v in self.arrs.items():
            setattr(c, k, _p(v))
        self.c = c
        if tvd:
            self.set_tvd(init_tvd(self, measure(self, mesh.node_xyz)))

    def set_inner(self, cell_inner):
        """Cell::is_inner per cell, for params area 1 / 2."""
        self.arrs["cell_inner"] = np.ascontiguousarray(cell_inner, np.uint8).reshape(-1)
        self.c.cell_inner = _p(self.arrs["cell_inner"])

    def set_tvd(self, tvd):
        self.arrs["face_tvd"] = np.ascontiguousarray(tvd, np.int32).reshape(-1, 2)
        self.c.face_tvd = _p(self.arrs["face_tvd"])


def params_c(params):
    p = OrParams()
    for k, _ in OrParams._fields_:
        setattr(p, k, params.get(k, PARAM_DEFAULTS[k]) if k in PARAM_DEFAULTS else params[k])
    return p


def measure(om, xyz, gcl=False):
    m = om.m
    out = dict(cell_center=np.empty((m.n_cells, 3)), cell_volume=np.empty(m.n_cells),
               face_center=np.empty((m.n_faces, 3)), face_normal=np.empty((m.n_faces, 3)), face_area=np.empty(m.n_faces))
    g = OrGeom(*[_p(out[k]) for k, _ in OrGeom._fields_])
    x = np.ascontiguousarray(xyz, np.float64)
    (lib().or_measure_gcl if gcl else lib().or_measure)(ctypes.byref(om.c), _p(x), ctypes.byref(g))
    return out


def face_flux(om, params, prim, geom, x1=None, x2=None, dt=1e-5, scal=None):
    """(rc, flux [nf,9], dt_face, branch), and with the passive scalars scal [nc,2] their fluxes
    kflux [nf,2] as a fifth item."""
    m = om.m
    g = OrGeom(*[_p(geom[k]) for k, _ in OrGeom._fields_])
    prim = np.ascontiguousarray(prim, np.float64)
    flux = np.empty((m.n_faces, 9))
    dtf = np.empty(m.n_faces)
    br = np.empty(m.n_faces, np.uint8)
    x1 = None if x1 is None else np.ascontiguousarray(x1, np.float64)
    x2 = None if x2 is None else np.ascontiguousarray(x2, np.float64)
    if scal is None:
        rc = lib().or_face_flux(ctypes.byref(om.c), ctypes.byref(params_c(params)), _p(prim), ctypes.byref(g), _p(x1),
                                _p(x2), float(dt), _p(flux), _p(dtf), _p(br))
        return rc, flux, dtf, br
    sc = np.ascontiguousarray(scal, np.float64).reshape(m.n_cells, 2)
    kflux = np.empty((m.n_faces, 2))
    rc = lib().or_face_flux_k(ctypes.byref(om.c), ctypes.byref(params_c(params)), _p(prim), _p(sc), ctypes.byref(g),
                              _p(x1), _p(x2), float(dt), _p(flux), _p(kflux), _p(dtf), _p(br))
    return rc, flux, dtf, br, kflux


def go(om, params, xyz, prim, steps, gcl=False, nthreads=1, disp=None, t_sim=0.0, h_fluids=None, fit=None,
       scal=None):
    """Runs or_go; returns (rc, prim_final, xyz_final, dt_hist, t_sim), and with the passive scalars
    scal [n_cells, 2] (Q, rho_He) their final values as a sixth item.  h_fluids [n_cells, n_h, 5]
    (rho_H, p_H, V_H per hydrogen fluid) adds the Calc_sourse_MF_Bera sources to every step; fit
    (OracleFit) with move_mesh 2 moves the nodes by the fitted surfaces and their rays."""
    m = om.m
    coords = np.zeros((2, m.n_nodes, 3))
    coords[0] = xyz
    state = np.zeros((2, m.n_cells, 8))
    state[0] = prim
    ks = None
    if scal is not None:
        ks = np.zeros((2, m.n_cells, 2))
        ks[0] = scal
    dth = np.zeros(max(1, steps))
    layer = ctypes.c_int(0)
    ts = ctypes.c_double(t_sim)
    d = None if disp is None else np.ascontiguousarray(disp, np.float64)
    hf = None if h_fluids is None else np.ascontiguousarray(h_fluids, np.float64).reshape(m.n_cells, -1, 5)
    rc = lib().or_go_k(ctypes.byref(om.c), ctypes.byref(params_c(params)), _p(coords), _p(state), _p(ks), _p(d),
                       int(steps), 1 if gcl else 0, int(nthreads), ctypes.byref(layer), _p(dth), ctypes.byref(ts),
                       0 if hf is None else hf.shape[1], _p(hf), None if fit is None else ctypes.byref(fit.c(params)))
    out = (rc, state[layer.value].copy(), coords[layer.value].copy(), dth[:steps], ts.value)
    return out if ks is None else out + (ks[layer.value].copy(),)


def chlld_batch(n_state, normal, qL, qR, w=None, n_disc=None, flags=None, kL=None, kR=None):
    qL = np.ascontiguousarray(qL, np.float64).reshape(-1, 8)
    qR = np.ascontiguousarray(qR, np.float64).reshape(-1, 8)
    n = qL.shape[0]
    ns = np.ascontiguousarray(np.broadcast_to(n_state, (n,)), np.int8)
    nd = None if n_disc is None else np.ascontiguousarray(np.broadcast_to(n_disc, (n,)), np.int8)
    fl = None if flags is None else np.ascontiguousarray(np.broadcast_to(flags, (n,)), np.uint8)
    nrm = np.ascontiguousarray(normal, np.float64).reshape(n, 3)
    ww = None if w is None else np.ascontiguousarray(np.broadcast_to(w, (n,)), np.float64)
    K = 0 if kL is None else int(np.asarray(kL).reshape(n, -1).shape[1])
    kLa = None if K == 0 else np.ascontiguousarray(kL, np.float64).reshape(n, K)
    kRa = None if K == 0 else np.ascontiguousarray(kR, np.float64).reshape(n, K)
    q = np.zeros((n, 8))
    k = np.zeros((n, max(K, 1)))
    ds = np.zeros((n, 3))
    br = np.zeros(n, np.uint8)
    rc = lib().or_chlld_batch(n, _p(ns), _p(nd), _p(fl), _p(nrm), _p(ww), _p(qL), _p(qR), K, _p(kLa), _p(kRa), _p(q),
                              _p(k) if K else None, _p(ds), _p(br))
    return q, k[:, :K], ds, br, rc


def get_bazis(n):
    n = np.ascontiguousarray(n, np.float64)
    t, m = np.zeros(3), np.zeros(3)
    lib().or_get_bazis(_p(n), _p(t), _p(m))
    return t, m


def get_potok(gamma, prim, n):
    out = np.zeros(8)
    lib().or_get_potok(gamma, prim[0], prim[4], prim[1], prim[2], prim[3], prim[5], prim[6], prim[7], n[0], n[1], n[2],
                       _p(out))
    return out


def new_connect(cell_nodes, n_nodes):
    cn = np.ascontiguousarray(cell_nodes, np.int32).reshape(-1)
    nc = cn.size // 8
    fn = np.empty(24 * nc, np.int32)
    fc = np.empty(12 * nc, np.int32)
    cf = np.empty(6 * nc, np.int32)
    nf = lib().or_new_connect(nc, _p(cn), int(n_nodes), _p(fn), _p(fc), _p(cf))
    return fn[:4 * nf].reshape(nf, 4), fc[:2 * nf].reshape(nf, 2), cf.reshape(nc, 6)


def cx_source(n_p, cp2, vp, n_h, ch2, vh, nodes=100):
    out = np.zeros(5)
    vp = np.ascontiguousarray(vp, np.float64)
    vh = np.ascontiguousarray(vh, np.float64)
    lib().or_cx_source(n_p, cp2, _p(vp), n_h, ch2, _p(vh), nodes, nodes, _p(out))
    return out


def mlp_forward(params, x):
    P = np.ascontiguousarray(params, np.float64)
    x = np.ascontiguousarray(x, np.float64).reshape(-1, 8)
    y = np.zeros((x.shape[0], 5))
    lib().or_mlp_forward(_p(P), x.shape[0], _p(x), _p(y))
    return y


def mlp_forward_bf16(params, x):
    """The network with the BF16 matrix-core inference's roundings (BF16 weights, features and
    activations; FP32 biases)."""
    P = np.ascontiguousarray(params, np.float64)
    x = np.ascontiguousarray(x, np.float64).reshape(-1, 8)
    y = np.zeros((x.shape[0], 5))
    lib().or_mlp_forward_bf16(_p(P), x.shape[0], _p(x), _p(y))
    return y


def mlp_loss_grad(params, x, y):
    P = np.ascontiguousarray(params, np.float64)
    x = np.ascontiguousarray(x, np.float64).reshape(-1, 8)
    y = np.ascontiguousarray(y, np.float64).reshape(-1, 5)
    g = np.zeros(P.size)
    loss = lib().or_mlp_loss_grad(_p(P), x.shape[0], _p(x), _p(y), _p(g))
    return loss, g


def face_velocities(om, x1, x2, geom1, dt, gcl=False):
    w = np.empty(om.m.n_faces)
    x1 = np.ascontiguousarray(x1, np.float64)
    x2 = np.ascontiguousarray(x2, np.float64)
    lib().or_face_velocities(ctypes.byref(om.c), _p(x1), _p(x2), _p(geom1["face_normal"]), _p(geom1["face_area"]),
                             float(dt), 1 if gcl else 0, _p(w))
    return w


def update(om, params, prim_old, vol_old, vol_new, flux, dt, src=None, scal=None, kflux=None):
    """(rc, prim_new, n_rho_floor, n_p_floor), and with the passive scalars scal [nc,2] and their
    face fluxes kflux [nf,2] the new scalars as a fifth item."""
    m = om.m
    out = np.empty((m.n_cells, 8))
    nr = ctypes.c_int64(0)
    npf = ctypes.c_int64(0)
    po = np.ascontiguousarray(prim_old, np.float64)
    vo = np.ascontiguousarray(vol_old, np.float64)
    vn = np.ascontiguousarray(vol_new, np.float64)
    fl = np.ascontiguousarray(flux, np.float64)
    sr = None if src is None else np.ascontiguousarray(src, np.float64)
    if scal is None:
        rc = lib().or_update(ctypes.byref(om.c), ctypes.byref(params_c(params)), _p(po), _p(vo), _p(vn), _p(fl),
                             _p(sr), float(dt), _p(out), ctypes.byref(nr), ctypes.byref(npf))
        return rc, out, nr.value, npf.value
    so = np.ascontiguousarray(scal, np.float64).reshape(m.n_cells, 2)
    kf = np.ascontiguousarray(kflux, np.float64).reshape(m.n_faces, 2)
    sn = np.empty((m.n_cells, 2))
    rc = lib().or_update_k(ctypes.byref(om.c), ctypes.byref(params_c(params)), _p(po), _p(so), _p(vo), _p(vn), _p(fl),
                           _p(kf), _p(sr), float(dt), _p(out), _p(sn), ctypes.byref(nr), ctypes.byref(npf))
    return rc, out, nr.value, npf.value, sn


def cx_rate_ref(U, c, a2, n1=100, n2=100):
    return lib().or_cx_rate_ref(float(U), float(c), float(a2), int(n1), int(n2))


def sigma_ls(g):
    return lib().or_sigma_ls(float(g))


def det_cos(x):
    return lib().or_det_cos(float(x))


def face_states(om, params, prim, geom, f, scal=None):
    """Left/right states of face f (Setka::Snos_on_Gran, first order + boundary); with the passive
    scalars scal [nc,2] also their left/right values (kL, kR)."""
    g = OrGeom(*[_p(geom[k]) for k, _ in OrGeom._fields_])
    prim = np.ascontiguousarray(prim, np.float64)
    qL, qR = np.empty(8), np.empty(8)
    if scal is None:
        rc = lib().or_face_states(ctypes.byref(om.c), ctypes.byref(params_c(params)), int(f), _p(prim),
                                  ctypes.byref(g), _p(qL), _p(qR))
        return rc, qL, qR
    sc = np.ascontiguousarray(scal, np.float64)
    kL, kR = np.empty(2), np.empty(2)
    rc = lib().or_face_states_k(ctypes.byref(om.c), ctypes.byref(params_c(params)), int(f), _p(prim), _p(sc),
                                ctypes.byref(g), _p(qL), _p(qR), _p(kL), _p(kR))
    return rc, qL, qR, kL, kR


def motion(x_old, disp, dt, t, omega, phase):
    x_old = np.ascontiguousarray(x_old, np.float64)
    disp = np.ascontiguousarray(disp, np.float64)
    out = np.empty_like(x_old)
    lib().or_motion(x_old.shape[0], _p(x_old), _p(disp), float(dt), float(t), float(omega), float(phase), _p(out))
    return out


def init_tvd(om, geom0):
    """TVD neighbours [n_faces, 2] (AA, BB; -1 = none) from the layer-0 geometry."""
    g = OrGeom(*[_p(geom0[k]) for k, _ in OrGeom._fields_])
    out = np.empty((om.m.n_faces, 2), np.int32)
    lib().or_init_tvd(ctypes.byref(om.c), ctypes.byref(g), _p(out))
    return out


def bera_source(prim, h_fluids, a2=0.102046, n_h_lism=3.0, kn=40.0906, scal=None, cell_type=None):
    """Setka::Calc_sourse_MF_Bera plasma sources [n, 5] = (rho, m_x, m_y, m_z, E) of states prim [n, 8]
    with hydrogen fluids h_fluids [n, n_h, 5]; with the passive scalars scal [n, 2] (Q, rho_He) and
    the cells' types cell_type [n] (1..4), determ_zone's proton split."""
    prim = np.ascontiguousarray(prim, np.float64).reshape(-1, 8)
    n = prim.shape[0]
    hf = np.ascontiguousarray(h_fluids, np.float64).reshape(n, -1, 5)
    out = np.empty((n, 5))
    L = lib()
    for c in range(n):
        if scal is None:
            L.or_bera_source(_p(prim[c]), hf.shape[1], _p(hf[c]), a2, n_h_lism, kn, _p(out[c]))
        else:
            L.or_bera_source_k(_p(prim[c]), float(scal[c][0]), float(scal[c][1]), int(cell_type[c]), hf.shape[1],
                               _p(hf[c]), a2, n_h_lism, kn, _p(out[c]))
    return out


def bera_source_h(prim, h_fluids, a2=0.102046, n_h_lism=3.0, kn=40.0906, scal=None, cell_type=None, arise=None):
    """or_bera_source_h per cell: (plasma sources [n, 5], hydrogen-fluid sources [n, n_h, 5]); arise
    [4, n_h] (1-based target fluid per zone and source fluid) or None for fluid = zone."""
    prim = np.ascontiguousarray(prim, np.float64).reshape(-1, 8)
    n = prim.shape[0]
    hf = np.ascontiguousarray(h_fluids, np.float64).reshape(n, -1, 5)
    nh = hf.shape[1]
    out = np.empty((n, 5))
    hout = np.empty((n, nh, 5))
    ar = None if arise is None else np.ascontiguousarray(arise, np.int8).reshape(4, nh)
    L = lib()
    for c in range(n):
        Q, He, ct = (0.0, 0.0, 1) if scal is None else (float(scal[c][0]), float(scal[c][1]), int(cell_type[c]))
        L.or_bera_source_h(_p(prim[c]), Q, He, ct, nh, _p(hf[c]), a2, n_h_lism, kn, _p(ar), _p(out[c]), _p(hout[c]))
    return out, hout


def determ_zone(Q, rho, cell_type):
    return lib().or_determ_zone(float(Q), float(rho), int(cell_type))


# ---- fitted surfaces ---------------------------------------------------------------------------
class OrFit(ctypes.Structure):
    _fields_ = [("move", ctypes.c_int32 * 4), ("sglag", ctypes.c_int32 * 4), ("velocity", _D * 4),
                ("k_smooth", _D * 4), ("phi_lim", _D), ("n_sn", _I64), ("sn_node", _P), ("sn_kind", _P),
                ("sn_face_off", _P), ("sn_face", _P), ("sn_nb_off", _P), ("sn_nb", _P), ("sn_centre", _P),
                ("n_rays", _I64), ("npr", ctypes.c_int32), ("ray_nodes", _P), ("ray_angles", _P),
                ("R0", _D), ("R1", _D), ("R5", _D), ("da1", _D), ("da2", _D), ("da3", _D), ("da4", _D), ("da5", _D),
                ("M0", ctypes.c_int32), ("M1", ctypes.c_int32), ("M11", ctypes.c_int32), ("M2", ctypes.c_int32),
                ("M3", ctypes.c_int32), ("M4", ctypes.c_int32), ("ext", _P), ("sn_cc", _P), ("sn_grid", _P),
                ("k_angle", _D), ("k_hp", _D), ("hp_angle", _D), ("hp_along", _D), ("n_tail", ctypes.c_int32),
                ("tail_len", ctypes.c_int32), ("N4", ctypes.c_int32), ("tail", _P), ("n_bs", ctypes.c_int32),
                ("bs_ref", _P), ("bs_off", _P), ("bs_nodes", _P)]

# Phys_param.cpp:513-519
HP_GRID_DEFAULTS = dict(sglag_HP_k_angle=0.001, sglag_HP_k=0.03, sglag_HP_angle=1.8, sglag_HP_along=1.0)


class OrRaysExt(ctypes.Structure):
    _fields_ = [("n", _I64), ("type", _P), ("off", _P), ("nodes", _P), ("opor", _P), ("angles", _P),
                ("R0", _D), ("R1", _D), ("R5", _D), ("L6", _D), ("L7", _D)] + \
               [(k, ctypes.c_int32) for k in ("M0", "M1", "M11", "M2", "M3", "M4", "MF", "N4", "N5")] + \
               [(k, _D) for k in ("dd1", "dd2", "dd3", "dd4", "dd5", "dd6", "dd7", "dd8", "ba1", "ba2", "ba3", "ba4",
                                  "ba5", "md1", "md2", "md3", "ea1", "ea2", "ea3")]


# Luch::type of the other ray families (Luch.cpp:153-739)
RAY_B, RAY_C, RAY_D, RAY_E, RAY_H, RAY_G = 1, 2, 3, 4, 5, 6
# Geo_param.h:40-84 defaults, with the values Setka.cpp:1130-1136 sets for every run
EXT_GEO_DEFAULTS = dict(dd1=1.0, dd2=1.0, dd3=19.0, dd4=15.0, dd5=19.0, dd6=14.0, dd7=0.7, dd8=1.0, ba1=1.0, ba2=1.0,
                        ba3=1.0, ba4=1.0, ba5=0.3, md1=0.2, md2=0.5, md3=0.3, ea1=0.1, ea2=0.5, ea3=0.2,
                        L6=-60.0, L7=-160.0)


class RaysExt:
    """The other ray families as flat arrays: type [n] int8, off [n+1], nodes, opor [n, 4] (-1 unused),
    angles [n, 2] (the, phi); geo holds R0, R1, R5, M0..M4, MF, N4, N5 and the EXT_GEO_DEFAULTS keys."""

    def __init__(self, types, node_lists, opor, angles, geo):
        self.type = np.ascontiguousarray(types, np.int8)
        self.off = np.zeros(len(node_lists) + 1, np.int64)
        self.off[1:] = np.cumsum([len(x) for x in node_lists])
        self.nodes = np.ascontiguousarray(np.concatenate([np.asarray(x, np.int32) for x in node_lists])
                                          if node_lists else np.zeros(0, np.int32), np.int32)
        self.opor = np.ascontiguousarray(opor, np.int32).reshape(-1, 4)
        self.angles = np.ascontiguousarray(angles, np.float64).reshape(-1, 2)
        self.geo = dict(EXT_GEO_DEFAULTS)
        self.geo.update(geo)

    def c(self):
        r = OrRaysExt()
        r.n = len(self.type)
        r.type, r.off, r.nodes, r.opor, r.angles = (_p(self.type), _p(self.off), _p(self.nodes), _p(self.opor),
                                                   _p(self.angles))
        for k, _ in OrRaysExt._fields_[6:]:
            setattr(r, k, self.geo[k])
        return r


def dvigenie_ext(ext, x):
    """or_dvigenie_ext on a copy of x."""
    xn = np.ascontiguousarray(x, np.float64).copy()
    lib().or_dvigenie_ext(ctypes.byref(ext.c()), _p(xn))
    return xn


# Phys_param.cpp:500-520
FIT_DEFAULTS = dict(move_TS=1, move_HP=1, move_BS=1, sglag_TS=1, sglag_HP=1, sglag_BS=0, velocity_TS=0.1,
                    velocity_HP=0.1, velocity_BS=0.3, sglag_TS_k_sphere=0.01, sglag_HP_k_sphere=0.01, sglag_BS_k=0.05)


def polar_angle(x, y):
    """Help_funk.cpp:78-107."""
    if abs(x) + abs(y) < 0.000001:
        return 0.0
    if x < 0:
        return math.atan(y / x) + math.pi
    if x > 0 and y >= 0:
        return math.atan(y / x)
    if x > 0 and y < 0:
        return math.atan(y / x) + 2.0 * math.pi
    if y > 0:
        return math.pi / 2.0
    if y < 0:
        return 3.0 * math.pi / 2.0
    return 0.0


def sphere_fit(points):
    """Smooth_head_TS3's centre (Setka_move.cpp:2257-2355): least squares of 2 p.c = |p|^2 (minimum
    norm, singular values below 3 eps of the largest dropped, as Eigen's SVD solve does), then up
    to 8 passes reweighted by Tukey biweights of |dist - mean dist| with cutoff 2.5 standard
    deviations, stopping when the rms residual changes by less than 1e-6."""
    P = np.asarray(points, np.float64).reshape(-1, 3)
    n = P.shape[0]
    A = 2.0 * P
    b = P[:, 0] * P[:, 0] + P[:, 1] * P[:, 1] + P[:, 2] * P[:, 2]
    rc = 3.0 * np.finfo(np.float64).eps
    c = np.linalg.lstsq(A, b, rcond=rc)[0]
    prev = np.finfo(np.float64).max
    for _ in range(8):
        d = np.linalg.norm(P - c, axis=1)
        radius = d.mean()
        res = np.abs(d - radius)
        rms = math.sqrt((res * res).sum() / n)
        if abs(prev - rms) < 1e-6:
            break
        prev = rms
        mean = res.mean()
        cutoff = 2.5 * math.sqrt(((res - mean) ** 2).sum() / n)
        with np.errstate(divide="ignore", invalid="ignore"):
            t = res / cutoff
            w = np.where(res < cutoff, (1.0 - t * t) * (1.0 - t * t), 0.0)
        c = np.linalg.lstsq(w[:, None] * A, w * b, rcond=rc)[0]
    return c


class OracleFit:
    """The fitted-surface set-up of a mesh in reference numbering, derived here from the face flags
    (bit3 TS, bit2 HP, bit4 BS): surface faces; HP faces with fewer than 4 surface neighbours
    (macros4, Setka.cpp:58-73: the same-surface faces of the cells beside cells[0]) move nothing;
    each surface node's faces in ascending id; its smoothing neighbours and sphere centre from
    node layer 0 (smooth(kind, x0): Smooth_head_TS3 / Smooth_head_HP3 / Find_Yzel_Sosed_for_BS);
    the rays (Luch::dvigenie data)."""

    def __init__(self, mesh, rays):
        fl = np.zeros(mesh.n_faces, np.uint8) if mesh.face_flags is None else np.asarray(mesh.face_flags, np.uint8)
        kind = np.where((fl >> 3) & 1, 1, np.where((fl >> 2) & 1, 2, np.where((fl >> 4) & 1, 3, 0))).astype(np.uint8)
        fn = np.asarray(mesh.face_nodes)
        fc = np.asarray(mesh.face_cells)
        kind[fc[:, 1] < 0] = 0  # only faces between two cells are fitted surfaces
        cf = np.asarray(mesh.cell_faces)
        sf = np.nonzero(kind)[0]
        moves = np.ones(mesh.n_faces, bool)
        for f in sf[kind[sf] == 2]:
            A = fc[f, 0]
            cnt = 0
            for g in cf[A]:
                if g == f:
                    continue
                B = fc[g, 1] if fc[g, 0] == A else fc[g, 0]
                if B < 0:
                    continue
                cnt += int(sum(1 for h in cf[B] if kind[h] == 2 and h != f))
            moves[f] = cnt >= 4
        nkind = np.zeros(mesh.n_nodes, np.uint8)
        faces_of = {}
        cand = {}
        for f in sf:  # ascending face id
            for v in fn[f]:
                nkind[v] = kind[f]
                if moves[f]:
                    faces_of.setdefault(int(v), []).append(int(f))
                cand.setdefault(int(v), set()).update(int(u) for u in fn[f])
        self.kind = kind
        self.sn = np.nonzero(nkind)[0].astype(np.int32)
        self.sn_kind = nkind[self.sn].astype(np.uint8)
        off = [0]
        lst = []
        for v in self.sn:
            lst += faces_of.get(int(v), [])
            off.append(len(lst))
        self.sn_face_off = np.asarray(off, np.int32)
        self.sn_face = np.asarray(lst, np.int32)
        self.cand = [sorted(cand[int(v)] | {int(v)}) for v in self.sn]
        self.nb = [[] for _ in self.sn]
        self.centre = np.zeros((len(self.sn), 3))
        self.rays = dict(rays)
        self.rays["ray_nodes"] = np.ascontiguousarray(rays["ray_nodes"], np.int32)
        self.rays["ray_angles"] = np.ascontiguousarray(rays["ray_angles"], np.float64)
        self.phi_lim = float(rays["phi_a"]) + math.pi / 180.0
        self._keep = None
        self.ext = None  # RaysExt of the other ray families, moved after the A rays
        self.cc = np.full((len(self.sn), 2), np.nan)  # Smooth_angle_HP circle centres
        self.grid = np.full((len(self.sn), 8), -1, np.int32)
        self.hp_k = dict(HP_GRID_DEFAULTS)
        self.tail = np.zeros((0, 0), np.int32)
        self.N4 = 0
        self.bs_ref = np.zeros(0, np.int32)
        self.bs_lists = []
        self._sn_index = {int(v): i for i, v in enumerate(self.sn)}

    def smooth_angle_hp(self, nodes, nb6, x0, centres=None):
        """Setka::Smooth_angle_HP (Setka_move.cpp:2663-2760): the circle centre of each HP node and
        its 6 azimuthal neighbours in the yz plane of x0, and the neighbours added to the node's
        smoothing set.  centres [n, 2] replace the fit (the library's own, for bit-level comparison)."""
        x0 = np.asarray(x0, np.float64)
        nb6 = np.asarray(nb6, np.int64).reshape(-1, 6)
        for k, v in enumerate(nodes):
            i = self._sn_index[int(v)]
            if centres is None:
                pts = np.ascontiguousarray(x0[[int(v)] + list(nb6[k])][:, 1:3])
                self.cc[i] = circle_fit(pts)[0]
            else:
                self.cc[i] = centres[k]
            self.nb[i] = sorted(set(self.nb[i]) | {int(u) for u in nb6[k]})

    def set_hp_grid(self, nodes, nb8, tail=None, N4=0, bs_ref=(), bs_lists=(), **coef):
        """The macros1 neighbours of HP grid nodes (AA1, AA11, AA3, AA33, AA2, AA22, AA4, AA44), the
        D-ray HP nodes per azimuth for the tail correction, the BS correction lists, and the
        sglag_HP_k_angle / _k / _angle / _along coefficients."""
        self.grid[:] = -1
        for k, v in enumerate(nodes):
            self.grid[self._sn_index[int(v)]] = np.asarray(nb8[k], np.int32)
        self.tail = np.ascontiguousarray(tail if tail is not None else np.zeros((0, 0)), np.int32)
        self.N4 = int(N4)
        self.bs_ref = np.ascontiguousarray(bs_ref, np.int32)
        self.bs_lists = [np.asarray(x, np.int32) for x in bs_lists]
        self.hp_k = dict(HP_GRID_DEFAULTS)
        self.hp_k.update(coef)

    def smooth(self, kind, x0):
        x0 = np.asarray(x0, np.float64)
        for i, v in enumerate(self.sn):
            if self.sn_kind[i] != kind:
                continue
            if kind == 2:
                a = x0[v]
                if polar_angle(a[0], math.sqrt(0.0 * 0.0 + a[1] * a[1] + a[2] * a[2])) > self.phi_lim:
                    continue
            pts = self.cand[i]
            if len(pts) < 3:
                continue
            self.nb[i] = [u for u in pts if u != v]
            if kind != 3:
                self.centre[i] = sphere_fit(x0[pts])

    def c(self, params):
        q = dict(FIT_DEFAULTS)
        q.update({k: params[k] for k in FIT_DEFAULTS if k in params})
        nb_off = np.zeros(len(self.sn) + 1, np.int32)
        nb_off[1:] = np.cumsum([len(x) for x in self.nb])
        nb = np.asarray([u for x in self.nb for u in x], np.int32)
        centre = np.ascontiguousarray(self.centre, np.float64)
        r = self.rays
        self._keep = (nb_off, nb, centre)
        f = OrFit()
        for k, (a, b, c) in enumerate([("move_TS", "sglag_TS", "TS"), ("move_HP", "sglag_HP", "HP"),
                                       ("move_BS", "sglag_BS", "BS")], start=1):
            f.move[k] = int(q[a])
            f.sglag[k] = int(q[b])
            f.velocity[k] = float(q["velocity_" + c])
        f.k_smooth[1] = q["sglag_TS_k_sphere"]
        f.k_smooth[2] = q["sglag_HP_k_sphere"]
        f.k_smooth[3] = q["sglag_BS_k"]
        f.phi_lim = self.phi_lim
        f.n_sn = len(self.sn)
        f.sn_node, f.sn_kind = _p(self.sn), _p(self.sn_kind)
        f.sn_face_off, f.sn_face = _p(self.sn_face_off), _p(self.sn_face)
        f.sn_nb_off, f.sn_nb, f.sn_centre = _p(nb_off), _p(nb) if nb.size else None, _p(centre)
        f.n_rays = r["ray_nodes"].shape[0]
        f.npr = r["ray_nodes"].shape[1]
        f.ray_nodes, f.ray_angles = _p(r["ray_nodes"]), _p(r["ray_angles"])
        for k in ("R0", "R1", "R5", "da1", "da2", "da3", "da4", "da5", "M0", "M1", "M11", "M2", "M3", "M4"):
            setattr(f, k, r[k])
        if self.ext is not None:
            self._ext_c = self.ext.c()
            f.ext = ctypes.cast(ctypes.pointer(self._ext_c), ctypes.c_void_p)
        cc = np.ascontiguousarray(self.cc, np.float64)
        grid = np.ascontiguousarray(self.grid, np.int32)
        bs_off = np.zeros(len(self.bs_lists) + 1, np.int32)
        bs_off[1:] = np.cumsum([len(x) for x in self.bs_lists])
        bs_nodes = np.ascontiguousarray(np.concatenate(self.bs_lists) if self.bs_lists else np.zeros(1), np.int32)
        self._keep2 = (cc, grid, bs_off, bs_nodes)
        f.sn_cc, f.sn_grid = _p(cc), _p(grid)
        f.k_angle, f.k_hp = self.hp_k["sglag_HP_k_angle"], self.hp_k["sglag_HP_k"]
        f.hp_angle, f.hp_along = self.hp_k["sglag_HP_angle"], self.hp_k["sglag_HP_along"]
        f.n_tail, f.tail_len, f.N4 = self.tail.shape[0], self.tail.shape[1] if self.tail.ndim == 2 else 0, self.N4
        f.tail = _p(self.tail) if self.tail.size else None
        f.n_bs, f.bs_ref, f.bs_off, f.bs_nodes = len(self.bs_lists), _p(self.bs_ref), _p(bs_off), _p(bs_nodes)
        return f


def culc_velocity_surface(om, params, fit, x, prim, time, gcl=False):
    """or_culc_velocity_surface on layer coordinates x; returns (rc, x_next, node_velocity [n_sn, 3])
    with x_next a copy of x whose surface nodes have moved."""
    x = np.ascontiguousarray(x, np.float64)
    g = measure(om, x, gcl)
    G = OrGeom(*[_p(g[k]) for k, _ in OrGeom._fields_])
    xn = x.copy()
    vel = np.zeros((len(fit.sn), 3))
    fc = fit.c(params)
    rc = lib().or_culc_velocity_surface(ctypes.byref(om.c), ctypes.byref(params_c(params)), ctypes.byref(fc), _p(x),
                                        _p(np.ascontiguousarray(prim, np.float64)), ctypes.byref(G), float(time),
                                        _p(xn), _p(vel))
    return rc, xn, vel


def hp_terms(A, P8, centre, ring, c_angle_term, c_along, c_angle, time):
    """or_hp_terms: (rc, V [5, 3]) -- the angle term, then the four macros1 terms of node A."""
    A = np.ascontiguousarray(A, np.float64)
    P8 = np.ascontiguousarray(P8, np.float64).reshape(8, 3)
    ring = np.ascontiguousarray(ring, np.float64).reshape(-1, 3)
    V = np.zeros((5, 3))
    rc = lib().or_hp_terms(_p(A), _p(P8), float(centre[0]), float(centre[1]), _p(ring), ring.shape[0],
                           float(c_angle_term), float(c_along), float(c_angle), float(time), _p(V))
    return rc, V


def circle_fit(points):
    """or_circle_fit: (centre [2], radius) of points [n, 2]."""
    P = np.ascontiguousarray(points, np.float64).reshape(-1, 2)
    c = np.zeros(2)
    r = lib().or_circle_fit(P.shape[0], _p(P), _p(c))
    return c, r


def dvigenie(fit, x, params=None):
    """or_dvigenie on a copy of x."""
    xn = np.ascontiguousarray(x, np.float64).copy()
    lib().or_dvigenie(ctypes.byref(fit.c(params or {})), _p(xn))
    return xn


# ---- Monte-Carlo neutral densities --------------------------------------------------------------
class OrMk(ctypes.Structure):
    _fields_ = [("n_in", _I64), ("face", _P), ("count", _P), ("mu", _P), ("seed", ctypes.c_uint64),
                ("v_inf", _D), ("a2", _D), ("kn", _D), ("max_steps", ctypes.c_int32), ("mistake", ctypes.c_int32),
                ("nthreads", ctypes.c_int32), ("R0", _D), ("zone", ctypes.c_int32), ("pad_", ctypes.c_int32),
                ("cell_mk", _P), ("face_mask", _P), ("in_sort", _P), ("in_side", _P), ("dist", _P), ("scal", _P),
                ("pui_nw", ctypes.c_int32), ("pad3_", ctypes.c_int32), ("pui_wr", _D), ("pui_arise", _P),
                ("pui_sm", _P), ("pui_sp", _P)]


# proton_arise_1..4 of the run with pickup ions (Phys_param.cpp:93-126), parent 0 (a thermal proton),
# rows sort 1..4: the pickup-ion fluid (1, 2; 0: none) an exchange of that atom feeds, zone by zone
PUI_ARISE = np.array([[0, 1, 1, 1], [2, 0, 2, 2], [1, 1, 0, 1], [1, 1, 1, 0]], np.int8)


def _pui(mk, m, pui):
    """Attaches the pickup-ion tallies of pui = dict(nw, wr, arise=PUI_ARISE) to mk; returns the
    raw arrays (sm [n_cells, nw], sp [n_cells, nw, 2]) or None."""
    if not pui:
        return None
    nw = int(pui["nw"])
    sm = np.zeros((m.n_cells, nw))
    sp = np.zeros((m.n_cells, nw, 2))
    ar = np.ascontiguousarray(pui.get("arise", PUI_ARISE), np.int8).reshape(16)
    mk.pui_nw, mk.pui_wr, mk.pui_arise, mk.pui_sm, mk.pui_sp = nw, float(pui["wr"]), _p(ar), _p(sm), _p(sp)
    return sm, sp, ar


def mk_pui_finish(vol, sm, sp, wr, a2, kn):
    """or_mk_pui_finish: (sm_out [n, nw], sp_out [n, nw, 2]) from raw tallies."""
    sm = np.ascontiguousarray(sm, np.float64)
    sp = np.ascontiguousarray(sp, np.float64)
    n, nw = sm.shape
    so, po = np.zeros_like(sm), np.zeros_like(sp)
    lib().or_mk_pui_finish(n, nw, float(wr), float(a2), float(kn), _p(np.ascontiguousarray(vol, np.float64)), _p(sm),
                           _p(sp), _p(so), _p(po))
    return so, po


class OrMkDist(ctypes.Structure):
    _fields_ = [("n_slots", ctypes.c_int32), ("r", ctypes.c_int32), ("slot", _P), ("w", _P)]


class MkDist:
    """The velocity distributions of the zone borders (or_mk_dist): a slot per interior face some
    zone lists, [slot, sort 1..4, side 0/1, 108 r^3 bins] of deposited weight."""

    def __init__(self, mesh, face_mask, r=1):
        fm = np.asarray(face_mask)
        listed = (fm != 0) & (np.asarray(mesh.face_cells)[:, 1] >= 0) & (np.asarray(mesh.face_bc) == 0)
        self.slot = np.full(mesh.n_faces, -1, np.int32)
        self.slot[listed] = np.arange(int(listed.sum()), dtype=np.int32)
        self.r = int(r)
        self.w = np.zeros((int(listed.sum()), 4, 2, 108 * self.r ** 3))
        self.c = OrMkDist(self.w.shape[0], self.r, _p(self.slot), _p(self.w))

    def bins(self):
        return 3 * self.r, 6 * self.r, 6 * self.r


MK_STATS = ("particles", "exchanges", "left", "sun_passes", "lost", "births_dropped", "segments",
            "exchange_capped", "zone_border")


def inner_radius(mesh, xyz):
    """Radius of the inner (Sun's) sphere: the largest distance from the centre of an inner boundary
    face's node (the shells' inner nodes lie on it); 0 without inner boundary faces."""
    f = np.flatnonzero(np.asarray(mesh.face_bc) == 1)
    if f.size == 0:
        return 0.0
    return float(np.linalg.norm(np.asarray(xyz)[np.asarray(mesh.face_nodes)[f].reshape(-1)], axis=1).max())


def mk_setup(om, geom, v_inf, n_per_face):
    """Injection faces, particles per face and weights (or_mk_setup)."""
    m = om.m
    face = np.zeros(m.n_faces, np.int32)
    count = np.zeros(m.n_faces, np.int32)
    mu = np.zeros(m.n_faces)
    g = OrGeom(*[_p(geom[k]) for k, _ in OrGeom._fields_])
    k = lib().or_mk_setup(ctypes.byref(om.c), ctypes.byref(g), float(v_inf), int(n_per_face), _p(face), _p(count),
                          _p(mu))
    return face[:k].copy(), count[:k].copy(), mu[:k].copy()


def mk_go(om, xyz, prim, v_inf, a2, kn, n_per_face, seed=1, max_steps=100000, geom=None, mistake=0, nthreads=1,
          scal=None, pui=None):
    """or_mk_go on the nodes xyz and state prim (geom: measure(om, xyz), computed when None;
    mistake: or_mk.mistake, nonzero only in tests that show the pins catch an error; scal: the
    passive scalars [n_cells, 2] = Q, rho_He, or None for rho_He = 0; pui: dict(nw, wr, arise) for
    the pickup-ion tallies, which then come back as a third item, the raw (sm, sp));
    returns (densities [n_cells, 5] = sum over sorts, sorts 1..4, per unit volume; stats dict)."""
    m = om.m
    x = np.ascontiguousarray(xyz, np.float64)
    geom = measure(om, x) if geom is None else geom
    face, count, mu = mk_setup(om, geom, v_inf, n_per_face)
    g = OrGeom(*[_p(geom[k]) for k, _ in OrGeom._fields_])
    mk = OrMk(len(face), _p(face), _p(count), _p(mu), int(seed), float(v_inf), float(a2), float(kn), int(max_steps),
              int(mistake), int(nthreads), inner_radius(m, x))
    sc = None if scal is None else np.ascontiguousarray(scal, np.float64).reshape(m.n_cells, 2)
    mk.scal = _p(sc)
    pt = _pui(mk, m, pui)
    tally = np.zeros((m.n_cells, 5))
    stats = np.zeros(8, np.int64)
    lib().or_mk_go(ctypes.byref(om.c), ctypes.byref(mk), _p(x), _p(np.ascontiguousarray(prim, np.float64)),
                   ctypes.byref(g), _p(tally), _p(stats))
    out = (tally / geom["cell_volume"][:, None], dict(zip(MK_STATS[:8], (int(v) for v in stats))))
    return out if pt is None else out + (pt[:2],)


def plasma_components_1(rho, p, rho_He, zone):
    """or_plasma_components_1: (rho_Th, p_Th, T_Th)."""
    th = (ctypes.c_double * 3)()
    lib().or_plasma_components_1(float(rho), float(p), float(rho_He), int(zone), th)
    return th[0], th[1], th[2]


def set_mk_zone(om, geom, mistake=0):
    """or_set_mk_zone: (rc, cell MK zone [n_cells], face zone mask [n_faces])."""
    m = om.m
    cell_mk = np.zeros(m.n_cells, np.int8)
    mask = np.zeros(m.n_faces, np.uint8)
    g = OrGeom(*[_p(geom[k]) for k, _ in OrGeom._fields_])
    rc = lib().or_set_mk_zone(ctypes.byref(om.c), ctypes.byref(g), int(mistake), _p(cell_mk), _p(mask))
    return rc, cell_mk, mask


def mk_setup_zone(om, geom, v_inf, n_per_face, zone, cell_mk, mask, dist):
    """or_mk_setup_zone: injections (face, sort, side, count, mu), or raises on rc < 0."""
    m = om.m
    n = 4 * m.n_faces
    face, count = np.zeros(n, np.int32), np.zeros(n, np.int32)
    sort, side, mu = np.zeros(n, np.int8), np.zeros(n, np.int8), np.zeros(n)
    g = OrGeom(*[_p(geom[k]) for k, _ in OrGeom._fields_])
    k = lib().or_mk_setup_zone(ctypes.byref(om.c), ctypes.byref(g), float(v_inf), int(n_per_face), int(zone),
                               _p(np.ascontiguousarray(cell_mk, np.int8)), _p(np.ascontiguousarray(mask, np.uint8)),
                               ctypes.byref(dist.c), _p(face), _p(sort), _p(side), _p(count), _p(mu))
    if k < 0:
        raise ValueError("or_mk_setup_zone: %d" % k)
    return face[:k].copy(), sort[:k].copy(), side[:k].copy(), count[:k].copy(), mu[:k].copy()


def mk_go_zone(om, xyz, prim, v_inf, a2, kn, n_per_face, zone, cell_mk, mask, dist, seed=1, max_steps=100000,
               geom=None, nthreads=1, scal=None, pui=None):
    """One pass of MK_go(zone) (or_mk_setup_zone + or_mk_go): the densities of the zone's cells per
    unit volume ([n_cells, 5], zero elsewhere), the stats dict, and the injections; deposits into
    dist, whose outgoing distributions of the zone's faces are cleared first (MK_go's Fill_null)."""
    m = om.m
    x = np.ascontiguousarray(xyz, np.float64)
    geom = measure(om, x) if geom is None else geom
    cm = np.ascontiguousarray(cell_mk, np.int8)
    mk_ = np.ascontiguousarray(mask, np.uint8)
    face, sort, side, count, mu = mk_setup_zone(om, geom, v_inf, n_per_face, zone, cm, mk_, dist)
    fc = np.asarray(m.face_cells)
    for f in np.flatnonzero((mk_ >> (zone - 1)) & 1):
        if dist.slot[f] >= 0:
            dist.w[dist.slot[f], :, 0 if cm[fc[f, 0]] == zone else 1, :] = 0.0
    g = OrGeom(*[_p(geom[k]) for k, _ in OrGeom._fields_])
    mk = OrMk(len(face), _p(face), _p(count), _p(mu), int(seed), float(v_inf), float(a2), float(kn), int(max_steps),
              0, int(nthreads), inner_radius(m, x), int(zone), 0, _p(cm), _p(mk_), _p(sort), _p(side),
              ctypes.cast(ctypes.pointer(dist.c), ctypes.c_void_p))
    sc = None if scal is None else np.ascontiguousarray(scal, np.float64).reshape(m.n_cells, 2)
    mk.scal = _p(sc)
    pt = _pui(mk, m, pui)
    tally = np.zeros((m.n_cells, 5))
    stats = np.zeros(9, np.int64)
    lib().or_mk_go(ctypes.byref(om.c), ctypes.byref(mk), _p(x), _p(np.ascontiguousarray(prim, np.float64)),
                   ctypes.byref(g), _p(tally), _p(stats))
    dens = tally / geom["cell_volume"][:, None]
    dens[cm != zone] = 0.0
    out = (dens, dict(zip(MK_STATS, (int(v) for v in stats))), (face, sort, side, count, mu))
    return out if pt is None else out + (pt[:2],)


def mk_deposit(dist, slot, sort, side, n, V, mu):
    lib().or_mk_deposit(ctypes.byref(dist.c), int(slot), int(sort), int(side),
                        _p(np.ascontiguousarray(n, np.float64)), _p(np.ascontiguousarray(V, np.float64)), float(mu))


def mk_sample(dist, slot, sort, side, n, n_draws, seed=1):
    V = np.zeros((int(n_draws), 3))
    lib().or_mk_sample_batch(ctypes.byref(dist.c), int(slot), int(sort), int(side),
                             _p(np.ascontiguousarray(n, np.float64)), int(n_draws), int(seed), _p(V))
    return V


def mk_change_velosity(n, U, V, cp, a2, seed=1, mistake=0):
    """n draws of the exchange sampler (or_mk_change_velosity_batch), units of cp: (W [n, 3], capped)."""
    W = np.zeros((n, 3))
    cap = ctypes.c_int64(0)
    u = np.ascontiguousarray(U, np.float64)
    v = np.ascontiguousarray(V, np.float64)
    lib().or_mk_change_velosity_batch(int(n), int(seed), float(a2), _p(u), _p(v), float(cp), int(mistake), _p(W),
                                      ctypes.byref(cap))
    return W, cap.value
