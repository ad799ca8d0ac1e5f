"""Synthetic inputs of the five workloads (meshes, boundary data, initial states, neutrals).

  cube   : 32^3 jittered hexahedra on [0,1]^3, Brio-Wu along x, HLLD everywhere, zero-gradient
           boundaries, CFL 0.4 (the reference CPU-runnable case, 100 steps).
  shell  : spherical shell 1..1500 AU, n_edge x n_edge cube-sphere quads per panel x n_layers
           log-spaced layers; inner solar-wind boundary, interstellar inflow (outer hard on the
           upwind half, outer soft elsewhere); HLLC on faces touching the r < 80 AU zone (Zone_1),
           HLLD elsewhere.  1M cells: n_edge 37 x 120 layers; 4M: 53 x 240.
  moving : the shell with prescribed node motion (radial, smooth in angle, tapered to zero at both
           boundaries, amplitude up to 5 % of the local radius).
Nondimensionalisation follows the reference's characteristic scales (Phys_param.h:117-122):
length 1 AU = 0.237455, density 0.06 cm^-3, velocity 10.3804 km/s, B 0.328842 uG
(= sqrt(char_p)), pressure 1.08137e-13 dyn/cm^2.
"""
import math
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import (BC_INNER_HARD, BC_INTERIOR, BC_OUTER_HARD, BC_OUTER_SOFT, BC_ZERO_GRADIENT, GEOM_REF,
               PREC_FP64, SRC_NONE, mesh_connect, mesh_cube, mesh_shell)

GAMMA = 5.0 / 3.0
R0_AU = 0.237455            # 1 AU in code units (Phys_param.h:117)
CHAR_RHO = 0.06             # cm^-3
CHAR_V = 10.3804            # km/s
CHAR_B_UG = 0.328842        # micro-gauss
CHAR_P = 1.08137e-13        # dyn/cm^2
K_B = 1.380649e-16
M_P = 1.67262192369e-24
V_INF = -2.54327            # Velosity_inf (Phys_param.h:95)
SPI4 = math.sqrt(4.0 * math.pi)


@dataclass
class Mesh:
    n_nodes: int
    n_cells: int
    n_faces: int
    cell_nodes: np.ndarray
    face_nodes: np.ndarray
    face_cells: np.ndarray
    cell_faces: np.ndarray
    face_bc: np.ndarray
    face_method: np.ndarray
    face_flags: Optional[np.ndarray]
    face_bstate: Optional[np.ndarray]
    cell_zone: np.ndarray
    cell_is_tvd: np.ndarray
    node_xyz: np.ndarray
    node_disp: Optional[np.ndarray] = None
    face_bscalar: Optional[np.ndarray] = None  # [n_faces, 2] Q, rho_He of the hard faces' fixed states


def splitmix_uniform(seed, index):
    """Counter-based U[0,1) (SplitMix64), identical to the C++ mesh generator's."""
    with np.errstate(over="ignore"):
        x = (np.uint64(seed) * np.uint64(0x2545F4914F6CDD1D) + np.asarray(index, dtype=np.uint64))
        x = x + np.uint64(0x9E3779B97F4A7C15)
        x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        x = x ^ (x >> np.uint64(31))
    return (x >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def default_params(**kw):
    p = dict(gamma=GAMMA, kfl=0.4, v_inf=V_INF, move_mesh=0, n_disc=0, geom_mode=GEOM_REF, precision=PREC_FP64,
             source_mode=SRC_NONE, cx_nodes=100, motion_omega=0.0, motion_phase=0.0, tvd=0, snos_on_hp=0)
    p.update(kw)
    return p


def _finish(xyz, cn, bc_fn, method_fn, zone, tvd_fn=None, bstate_fn=None, disp=None):
    fn, fcells, cf = mesh_connect(cn, xyz.shape[0])
    nf = fn.shape[0]
    fcen = xyz[fn].mean(axis=1)
    ccen = xyz[cn].mean(axis=1)
    boundary = fcells[:, 1] < 0
    bc = bc_fn(fcen, ccen, fn, fcells, boundary).astype(np.int8)
    bc[~boundary] = BC_INTERIOR
    method = method_fn(fcells, zone).astype(np.int8)
    bstate = bstate_fn(fcen, bc) if bstate_fn else None
    tvd = np.ones(cn.shape[0], np.uint8)
    if tvd_fn is not None:
        tvd = tvd_fn(fcells, bc).astype(np.uint8)
    return Mesh(n_nodes=xyz.shape[0], n_cells=cn.shape[0], n_faces=nf, cell_nodes=cn, face_nodes=fn,
                face_cells=fcells, cell_faces=cf, face_bc=bc, face_method=method, face_flags=None,
                face_bstate=bstate, cell_zone=zone.astype(np.int8), cell_is_tvd=tvd, node_xyz=xyz,
                node_disp=disp), ccen


# ---------------------------------------------------------------------------------------------
def cube_brio_wu(n=32, jitter=0.15, seed_mesh=1, seed_state=2):
    """Brio-Wu shock tube on a jittered n^3 cube; returns (mesh, prim0 [nc,8], params)."""
    xyz, cn = mesh_cube(n, jitter, seed_mesh)
    zone = np.full(cn.shape[0], 3, np.int8)
    mesh, cc = _finish(xyz, cn, lambda fc, cc, fn, fcl, b: np.where(b, BC_ZERO_GRADIENT, BC_INTERIOR),
                       lambda fcl, z: np.full(fcl.shape[0], 3), zone)
    nc = cn.shape[0]
    prim = np.zeros((nc, 8))
    left = cc[:, 0] < 0.5
    prim[:, 0] = np.where(left, 1.0, 0.125)
    prim[:, 4] = np.where(left, 1.0, 0.1)
    prim[:, 5] = 0.75 * SPI4
    prim[:, 6] = np.where(left, 1.0, -1.0) * SPI4
    noise = 2.0 * splitmix_uniform(seed_state, np.arange(nc)) - 1.0
    prim[:, 0] *= 1.0 + 0.01 * noise
    return mesh, prim, default_params()


# ---------------------------------------------------------------------------------------------
def solar_wind_state(x):
    """Supersonic wind: n=7 cm^-3, V=375 km/s radial, T=5e4 K at 1 AU; rho ~ r^-2,
    p ~ r^-2gamma; Parker B (|B| = 5 nT at 1 AU on the equator), B_r = -B0 (R0/r)^2,
    B_phi = -B_r sin(theta) r/R0 (the form of Setka_physics.cpp:510-513)."""
    x = np.asarray(x, dtype=np.float64)
    r = np.linalg.norm(x, axis=1)
    rho1 = 7.0 / CHAR_RHO
    v1 = 375.0 / CHAR_V
    p1 = 2.0 * 7.0 * K_B * 5.0e4 / CHAR_P
    b_au = 50.0 / CHAR_B_UG
    B0 = b_au / math.sqrt(2.0)
    s = R0_AU / r
    out = np.zeros((x.shape[0], 8))
    out[:, 0] = rho1 * s ** 2
    out[:, 1:4] = v1 * x / r[:, None]
    out[:, 4] = p1 * s ** (2.0 * GAMMA)
    theta = np.arccos(np.clip(x[:, 2] / r, -1.0, 1.0))
    phi = np.arctan2(x[:, 1], x[:, 0])
    br = -B0 * s ** 2
    bphi = -br * np.sin(theta) * (r / R0_AU)
    rhat = x / r[:, None]
    phat = np.stack([-np.sin(phi), np.cos(phi), np.zeros_like(phi)], axis=1)
    out[:, 5:8] = br[:, None] * rhat + bphi[:, None] * phat
    return out


def lism_state():
    """Interstellar inflow: n_p = 0.04 cm^-3, V = 26.4 km/s along -x, T = 7500 K, |B| = 4 uG at
    60 degrees to the flow direction."""
    rho = 0.04 / CHAR_RHO
    v = -26.4 / CHAR_V
    p = 2.0 * 0.04 * K_B * 7500.0 / CHAR_P
    b = 4.0 / CHAR_B_UG
    a = math.radians(60.0)
    return np.array([rho, v, 0.0, 0.0, p, -b * math.cos(a), b * math.sin(a), 0.0])


HE_FRACTION = 0.15  # rho_He / rho of the test states


def passive_scalars(mesh, prim, r_blend_au=100.0):
    """Initial passive scalars [n_cells, 2] of a shell state: Q/rho = 1 in the solar wind and 100 in
    the interstellar medium, blended over radius as the shell's state is (so Q/rho crosses the
    zone switch 50 of determ_zone), and rho_He = HE_FRACTION * rho."""
    ccen = mesh.node_xyz[mesh.cell_nodes].mean(axis=1)
    rc = np.linalg.norm(ccen, axis=1)
    w = 0.5 * (1.0 + np.tanh((rc - r_blend_au * R0_AU) / (0.2 * r_blend_au * R0_AU)))
    out = np.empty((mesh.n_cells, 2))
    out[:, 0] = (1.0 + 99.0 * w) * prim[:, 0]
    out[:, 1] = HE_FRACTION * prim[:, 0]
    return out


def _hard_face_scalars(mesh):
    """Passive scalars of the hard faces' fixed states [n_faces, 2]: Q/rho = 1 in the wind (inner
    sphere), 100 in the interstellar inflow (the reference's initial states, Setka_physics.cpp:
    272,542), rho_He = HE_FRACTION * rho."""
    bq = np.zeros((mesh.n_faces, 2))
    rho_b = mesh.face_bstate[:, 0]
    bq[:, 0] = np.where(mesh.face_bc == BC_OUTER_HARD, 100.0, 1.0) * rho_b
    bq[:, 1] = HE_FRACTION * rho_b
    return bq


def shell_radii(n_layers, r_out_au=1500.0):
    return R0_AU * r_out_au ** (np.arange(n_layers + 1) / n_layers)


def shell(n_edge=37, n_layers=120, moving=False, seed_motion=3, r_zone1_au=80.0, r_blend_au=100.0,
          amp=0.05, omega=100.0, noise=0.0, seed_state=4):
    """Heliosphere-like shell; returns (mesh, prim0, params).

    `noise` > 0 perturbs density and velocity cell by cell (relative amplitude `noise`, seeded).
    The unperturbed wind is exactly radial, so every lateral face has v.n = 0 up to rounding and
    its HLLC/HLLD contact speed SM sits on the branch point of the flux; the perturbation moves
    parity workloads off that set.
    """
    radii = shell_radii(n_layers)
    xyz, cn = mesh_shell(n_edge, radii)
    rin, rout = radii[0], radii[-1]
    rmid = math.sqrt(rin * rout)
    ccen = xyz[cn].mean(axis=1)
    rc = np.linalg.norm(ccen, axis=1)
    zone = np.where(rc < r_zone1_au * R0_AU, 1, 3).astype(np.int8)
    lism = lism_state()

    def bc_fn(fc, cc, fn, fcl, b):
        rf = np.linalg.norm(fc, axis=1)
        inner = b & (rf < rmid)
        outer = b & ~inner
        # outward normal at the outer sphere is radial: inflow where V_inf . n < 0, i.e. x > 0
        bc = np.full(fc.shape[0], BC_INTERIOR)
        bc[inner] = BC_INNER_HARD
        bc[outer & (fc[:, 0] > 0.0)] = BC_OUTER_HARD
        bc[outer & (fc[:, 0] <= 0.0)] = BC_OUTER_SOFT
        return bc

    def method_fn(fcl, z):
        za = z[fcl[:, 0]]
        zb = np.where(fcl[:, 1] >= 0, z[np.maximum(fcl[:, 1], 0)], 3)
        return np.where((za == 1) | (zb == 1), 2, 3)

    def bstate_fn(fc, bc):
        bs = np.zeros((fc.shape[0], 8))
        ih = bc == BC_INNER_HARD
        bs[ih] = solar_wind_state(fc[ih])
        bs[bc == BC_OUTER_HARD] = lism
        return bs

    def tvd_fn(fcl, bc):
        t = np.ones(cn.shape[0], np.uint8)
        t[fcl[bc == BC_INNER_HARD, 0]] = 0   # Init_physics: cells on Inner_Hard faces (Setka_physics.cpp:585)
        return t

    disp = None
    if moving:
        rn = np.linalg.norm(xyz, axis=1)
        th = np.arccos(np.clip(xyz[:, 2] / rn, -1.0, 1.0))
        ph = np.arctan2(xyz[:, 1], xyz[:, 0])
        u = splitmix_uniform(seed_motion, np.arange(3))
        phase = 2.0 * math.pi * u
        taper = np.clip((1.0 - rin / rn) * (1.0 - rn / rout) * 4.0, 0.0, 1.0)
        amp_n = amp * rn * np.sin(3.0 * th + phase[0]) * np.sin(2.0 * ph + phase[1]) * taper
        disp = amp_n[:, None] * xyz / rn[:, None]
    mesh, _ = _finish(xyz, cn, bc_fn, method_fn, zone, tvd_fn=tvd_fn, bstate_fn=bstate_fn, disp=disp)
    mesh.face_bscalar = _hard_face_scalars(mesh)
    # interior: r^-2 wind blended to the interstellar state around r_blend (tanh in r)
    wind = solar_wind_state(ccen)
    w = 0.5 * (1.0 + np.tanh((rc - r_blend_au * R0_AU) / (0.2 * r_blend_au * R0_AU)))
    prim = (1.0 - w)[:, None] * wind + w[:, None] * lism[None, :]
    if noise > 0.0:
        nc = cn.shape[0]
        u = [2.0 * splitmix_uniform(seed_state, np.arange(nc) * 4 + k) - 1.0 for k in range(4)]
        vmag = np.linalg.norm(prim[:, 1:4], axis=1)
        prim[:, 0] *= 1.0 + noise * u[0]
        for k in range(3):
            prim[:, 1 + k] += noise * vmag * u[1 + k]
    params = default_params()
    if moving:
        params.update(move_mesh=1, motion_omega=omega, motion_phase=float(2.0 * math.pi * splitmix_uniform(seed_motion, 3)))
    return mesh, prim, params


def sun_shell(n_edge=6, n_layers=8, r_out=4.0, v_inf=V_INF):
    """A small spherical shell around the Sun's sphere (radius 1, outer radius r_out, log-spaced
    layers), filled with the interstellar Maxwellian of the Monte-Carlo neutrals (rho_Th = 1,
    V = (v_inf, 0, 0), p_Th = 1/2: cp = 1): every outer face is an inflow face (bc 3), the inner
    sphere a hard face (bc 1).  For the neutrals' pass through the Sun's sphere, where the heliosphere
    shell's 1:1500 radii leave too few particles near it.  Returns (mesh, prim, params)."""
    radii = r_out ** (np.arange(n_layers + 1) / n_layers)
    xyz, cn = mesh_shell(n_edge, radii)
    zone = np.full(cn.shape[0], 3, np.int8)

    def bc_fn(fc, cc, fn, fcl, b):
        rf = np.linalg.norm(fc, axis=1)
        bc = np.full(fc.shape[0], BC_INTERIOR)
        bc[b & (rf < math.sqrt(r_out))] = BC_INNER_HARD
        bc[b & (rf >= math.sqrt(r_out))] = BC_OUTER_SOFT
        return bc

    me = 1.0 / 1836.15
    lism = np.array([1.0 + me, v_inf, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0])

    def bstate_fn(fc, bc):
        return np.tile(lism, (fc.shape[0], 1))

    mesh, _ = _finish(xyz, cn, bc_fn, lambda fcl, z: np.full(fcl.shape[0], 3), zone, bstate_fn=bstate_fn)
    prim = np.tile(lism, (mesh.n_cells, 1))
    return mesh, prim, default_params(v_inf=v_inf)


# ---------------------------------------------------------------------------------------------
def a_luch_radii(R0, R1, R2, R3, R4, R5, M0, M1, M11, M2, M3, M4, da1, da2, da3, da4, da5):
    """Radii [n_rays, npr] of the nodes of rays whose termination shock, heliopause and bow shock
    sit at R2, R3, R4 [n_rays]: the distribution of Luch::dvigenie for A_Luch rays (Luch.cpp:8-151),
    used here to lay out the initial mesh."""
    R2, R3, R4 = (np.asarray(a, np.float64).reshape(-1, 1) for a in (R2, R3, R4))
    cols = []
    for j in range(M0):
        cols.append(np.full_like(R2, R0 + j * (R1 - R0) / M0))
    cols.append(np.full_like(R2, R1))
    for j in range(M1):
        cols.append(R1 + (j + 1) * (R2 - R1) / (M1 + 1))
    cols.append(R2)
    dr = (R2 - R1) / (M1 + 1)
    for j in range(M11):
        cols.append(R2 + dr * (j + 1))
    a, b = da1, da2
    for j in range(M2 - M11):
        t = (j + 1.0) / (M2 - M11 + 1.0)
        u = 1.0 - 3.0 * t + 2.0 * t * t
        w = 1.0 - 2.0 * t
        ss = t * (a * u * u + t * (7.0 + b * w * w * (-1.0 + t) - 2.0 * t * (9.0 + 2.0 * t * (-5.0 + 2.0 * t))))
        cols.append(R2 + dr * M11 + (R3 - R2 - dr * M11) * ss)
    cols.append(R3)
    a, b = da3, da4
    for j in range(M3):
        t = (j + 1.0) / (M3 + 1.0)
        cols.append(R3 + (R4 - R3) * (t * (a - (2.0 * a + b - 3.0) * t + (a + b - 2.0) * t * t)))
    cols.append(R4)
    a, b = da5, 2.0
    for j in range(M4):
        t = (j + 1.0) / (M4 + 1.0)
        cols.append(R4 + (R5 - R4) * (a * t - (3.0 * a + b - 4.0) * t ** 3 + (2.0 * a + b - 3.0) * t ** 4))
    cols.append(np.full_like(R2, R5))
    return np.concatenate(cols, axis=1)


def shell_fitted(n_edge=6, M=(3, 6, 2, 8, 5, 5), R0_au=1.0, R1_au=10.0, R5_au=600.0, surf_au=(80.0, 120.0, 240.0),
                 asym=(0.2, 0.25, 0.3), da=(1.0, 1.0, 1.0, 1.0, 0.5), zone1=True, noise=0.0, seed_state=4, tvd=False,
                 bs_upwind=False):
    """Heliosphere shell whose nodes lie on A_Luch rays through the cube-sphere directions, with
    the termination shock, heliopause and bow shock as fitted node surfaces: the TS at
    surf_au[0] * (1 - asym[0] cos the) along a ray at polar angle `the` from the +x axis (closest
    upwind, where the interstellar flow comes from), the HP and the BS likewise; the ray layout of
    Luch::dvigenie with counts M = (M0, M1, M11, M2, M3, M4) and shapes da = (da1 .. da5).  Zones:
    1 inside the TS (2 when zone1 is False), 2 to the HP, 3 to the BS, 4 outside.  TS, HP (with
    the contact flag of contact_hard, Phys_param.cpp:523) and BS faces use HLLD (Gran::Get_method);
    other faces HLLC where they touch Zone_1, HLLD elsewhere.  bs_upwind: the bow shock is a fitted
    surface only where its faces' centres have x > 0, as on the reference's own mesh, whose tail has
    no bow shock (the Zone_3 / Zone_4 faces there are plain faces, Setka_MK.cpp:560-641).  Returns
    (mesh, prim, params, rays) with move_mesh = 2 and the surface-fitting defaults of
    Phys_param.cpp:500-520."""
    M0, M1, M11, M2, M3, M4 = M
    npr = M0 + M1 + M2 + M3 + M4 + 5
    nl = npr - 1
    R0, R1, R5 = R0_au * R0_AU, R1_au * R0_AU, R5_au * R0_AU
    xyz, cn = mesh_shell(n_edge, np.linspace(R0, R5, nl + 1))
    ns = 6 * n_edge * n_edge + 2
    unit = xyz[:ns] / np.linalg.norm(xyz[:ns], axis=1)[:, None]
    the = np.arccos(np.clip(unit[:, 0], -1.0, 1.0))
    phi = np.arctan2(unit[:, 2], unit[:, 1])
    ct = np.cos(the)
    Rs = [surf_au[k] * R0_AU * (1.0 - asym[k] * ct) for k in range(3)]
    rad = a_luch_radii(R0, R1, Rs[0], Rs[1], Rs[2], R5, M0, M1, M11, M2, M3, M4, *da)  # [ns, npr]
    for k in range(npr):
        r = rad[:, k]
        xyz[k * ns:(k + 1) * ns, 0] = r * np.cos(the)
        xyz[k * ns:(k + 1) * ns, 1] = r * np.sin(the) * np.cos(phi)
        xyz[k * ns:(k + 1) * ns, 2] = r * np.sin(the) * np.sin(phi)
    iTS, iHP, iBS = M0 + M1 + 1, M0 + M1 + M2 + 2, M0 + M1 + M2 + M3 + 3
    layer = cn.min(axis=1) // ns
    zone = np.where(layer < iTS, 1 if zone1 else 2, np.where(layer < iHP, 2, np.where(layer < iBS, 3, 4))).astype(np.int8)

    def bc_fn(fc, cc, fn, fcl, b):
        fl = fn.min(axis=1) // ns
        bc = np.full(fc.shape[0], BC_INTERIOR)
        bc[b & (fl == 0)] = BC_INNER_HARD
        outer = b & (fl > 0)
        bc[outer & (fc[:, 0] > 0.0)] = BC_OUTER_HARD
        bc[outer & (fc[:, 0] <= 0.0)] = BC_OUTER_SOFT
        return bc

    def surf_kind(fn):
        lo, hi = fn.min(axis=1) // ns, fn.max(axis=1) // ns
        flat = lo == hi
        return np.where(flat & (lo == iTS), 1, np.where(flat & (lo == iHP), 2, np.where(flat & (lo == iBS), 3, 0)))

    def method_fn(fcl, z):
        za = z[fcl[:, 0]]
        zb = np.where(fcl[:, 1] >= 0, z[np.maximum(fcl[:, 1], 0)], 3)
        return np.where((za == 1) | (zb == 1), 2, 3)

    def bstate_fn(fc, bc):
        bs = np.zeros((fc.shape[0], 8))
        ih = bc == BC_INNER_HARD
        bs[ih] = solar_wind_state(fc[ih])
        bs[bc == BC_OUTER_HARD] = lism_state()
        return bs

    mesh, ccen = _finish(xyz, cn, bc_fn, method_fn, zone, bstate_fn=bstate_fn)
    mesh.face_bscalar = _hard_face_scalars(mesh)
    sk = surf_kind(mesh.face_nodes)
    if bs_upwind:
        sk[(sk == 3) & (xyz[mesh.face_nodes].mean(axis=1)[:, 0] <= 0.0)] = 0
    flags = np.zeros(mesh.n_faces, np.uint8)
    flags[sk == 1] = 1 << 3
    flags[sk == 2] = (1 << 2) | (1 << 1)
    flags[sk == 3] = 1 << 4
    mesh.face_flags = flags
    mesh.face_method = np.where(sk > 0, 3, mesh.face_method).astype(np.int8)
    if tvd:
        t = np.ones(cn.shape[0], np.uint8)
        t[mesh.face_cells[mesh.face_bc == BC_INNER_HARD, 0]] = 0
        mesh.cell_is_tvd = t
    rc = np.linalg.norm(ccen, axis=1)
    wind = solar_wind_state(ccen)
    rhp = surf_au[1] * R0_AU
    w = 0.5 * (1.0 + np.tanh((rc - rhp) / (0.2 * rhp)))
    prim = (1.0 - w)[:, None] * wind + w[:, None] * lism_state()[None, :]
    if noise > 0.0:
        nc = cn.shape[0]
        u = [2.0 * splitmix_uniform(seed_state, np.arange(nc) * 4 + k) - 1.0 for k in range(4)]
        vmag = np.linalg.norm(prim[:, 1:4], axis=1)
        prim[:, 0] *= 1.0 + noise * u[0]
        for k in range(3):
            prim[:, 1 + k] += noise * vmag * u[1 + k]
    params = default_params(move_mesh=2, tvd=1 if tvd else 0, move_TS=1, move_HP=1, move_BS=1, sglag_TS=1,
                            sglag_HP=1, sglag_BS=0, velocity_TS=0.1, velocity_HP=0.1, velocity_BS=0.3,
                            sglag_TS_k_sphere=0.01, sglag_HP_k_sphere=0.01, sglag_BS_k=0.05)
    rays = dict(ray_nodes=(np.arange(npr)[None, :] * ns + np.arange(ns)[:, None]).astype(np.int32),
                ray_angles=np.stack([the, phi], axis=1), R0=R0, R1=R1, R5=R5, M0=M0, M1=M1, M11=M11, M2=M2,
                M3=M3, M4=M4, da1=da[0], da2=da[1], da3=da[2], da4=da[3], da5=da[4], phi_a=math.pi - math.pi / 180.0)
    return mesh, prim, params, rays


def inner_cells_from_rays(mesh, rays):
    """Cell::is_inner of Setka::Init_boundary_grans (Setka_physics.cpp:113-156): the first M0 + 8
    nodes of every ray are inner, and a cell is inner when all eight of its nodes are."""
    inner = np.zeros(mesh.n_nodes, bool)
    inner[np.asarray(rays["ray_nodes"])[:, :rays["M0"] + 1 + 7].reshape(-1)] = True
    return inner[mesh.cell_nodes].all(axis=1).astype(np.uint8)


def neutrals(n_cells, seed=5):
    """Per-cell neutral-H moments [n_H, c_H^2, V_H(3)] in code units: n_H ~ U(0.08, 0.2) cm^-3,
    V_H = 26 km/s along the inflow axis (-x) with +-30 % per-component perturbation, T_H ~
    U(6e3, 2e4) K (c_H^2 = 2 k T / m_p)."""
    idx = np.arange(n_cells, dtype=np.uint64) * np.uint64(6)
    u = [splitmix_uniform(seed, idx + np.uint64(k)) for k in range(6)]
    out = np.zeros((n_cells, 5))
    out[:, 0] = (0.08 + 0.12 * u[0]) / CHAR_RHO
    T = 6.0e3 + 1.4e4 * u[1]
    out[:, 1] = 2.0 * K_B * T / M_P / (CHAR_V * 1e5) ** 2
    v0 = 26.0 / CHAR_V
    out[:, 2] = -v0 + 0.3 * v0 * (2.0 * u[2] - 1.0)
    out[:, 3] = 0.3 * v0 * (2.0 * u[3] - 1.0)
    out[:, 4] = 0.3 * v0 * (2.0 * u[4] - 1.0)
    return out


def h_fluids(n_cells, n_h=4, seed=9):
    """Per-cell hydrogen fluids [n_cells, n_h, 5] = (rho_Hk, p_Hk, V_Hk) in code units, after the four
    populations of TITAN's multi-fluid hydrogen (pristine LISM atoms, and atoms born in the outer
    heliosheath, the inner heliosheath and the supersonic wind): density, temperature and speed
    along -x per population, with +-30 % seeded per-cell perturbations.  p_Hk = rho_Hk T_Hk / 2,
    the inverse of T_H = 2 p_H / rho_H (Setka_physics.cpp:822)."""
    pops = [(0.10, 6.5e3, 26.0), (0.05, 1.6e4, 18.0), (0.01, 3.0e5, 100.0), (0.004, 1.0e5, 400.0)]
    out = np.zeros((n_cells, n_h, 5))
    idx = np.arange(n_cells, dtype=np.uint64) * np.uint64(5 * n_h)
    for k in range(n_h):
        n0, T0, v0 = pops[k % len(pops)]
        u = [splitmix_uniform(seed, idx + np.uint64(5 * k + j)) for j in range(5)]
        rho = n0 / CHAR_RHO * (0.7 + 0.6 * u[0])
        T = 2.0 * K_B * (T0 * (0.7 + 0.6 * u[1])) / M_P / (CHAR_V * 1e5) ** 2
        v = v0 / CHAR_V
        out[:, k, 0] = rho
        out[:, k, 1] = 0.5 * rho * T
        out[:, k, 2] = -v * (0.7 + 0.6 * u[2])
        out[:, k, 3] = 0.3 * v * (2.0 * u[3] - 1.0)
        out[:, k, 4] = 0.3 * v * (2.0 * u[4] - 1.0)
    return out


WORKLOADS = {
    "cube32": dict(fn="cube", n=32),
    "shell1m": dict(fn="shell", n_edge=37, n_layers=120),
    "shell1m_moving": dict(fn="shell", n_edge=37, n_layers=120, moving=True),
    "shell4m_moving": dict(fn="shell", n_edge=53, n_layers=240, moving=True),
    # 240 cell layers on 6*53^2 rays: TS, HP and BS fitted (move_mesh 2)
    "shell4m_fitted": dict(fn="shell_fitted", n_edge=53, M=(10, 70, 6, 70, 50, 36), R5_au=1500.0),
}


def build(name, **over):
    spec = dict(WORKLOADS[name])
    spec.update(over)
    fn = spec.pop("fn")
    if fn == "cube":
        return cube_brio_wu(**spec)
    if fn == "shell_fitted":
        return shell_fitted(**spec)  # (mesh, prim, params, rays)
    return shell(**spec)
