"""Shared generators for the test suite."""
import math

import numpy as np


def random_states(rng, n, bmag=2.0, vmag=3.0):
    """Admissible primitive states (rho, V, p, B) in the reference's Gaussian units."""
    q = np.zeros((n, 8))
    q[:, 0] = np.exp(rng.uniform(np.log(0.05), np.log(20.0), n))
    q[:, 1:4] = rng.normal(0.0, vmag, (n, 3))
    q[:, 4] = np.exp(rng.uniform(np.log(0.01), np.log(50.0), n))
    q[:, 5:8] = rng.normal(0.0, bmag, (n, 3)) * np.sqrt(4.0 * np.pi)
    return q


def random_normals(rng, n):
    v = rng.normal(size=(n, 3))
    return v / np.linalg.norm(v, axis=1)[:, None]


def random_rotation(rng):
    a = rng.normal(size=(3, 3))
    q, r = np.linalg.qr(a)
    q *= np.sign(np.diag(r))[None, :]
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


def rel_err(a, b, floor=1e-300):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.abs(a - b) / np.maximum(np.maximum(np.abs(a), np.abs(b)), floor)


def var_err(a, b):
    """Error per variable relative to that variable's largest magnitude on the mesh.  A coarse
    norm: used only to show that a change is visible, never as a parity bound."""
    return (np.abs(a - b).max(axis=0) / np.maximum(np.abs(b).max(axis=0), 1e-300)).max()


GAMMA = 5.0 / 3.0


def conserved_state(prim):
    """Conserved variables per cell (rho, rho v, E, B) of primitive states (rho, v, p, B), B in
    the reference's Gaussian units: E = p/(gamma-1) + rho v^2/2 + B^2/8pi."""
    q = np.asarray(prim, np.float64).reshape(-1, 8)
    U = np.empty_like(q)
    U[:, 0] = q[:, 0]
    U[:, 1:4] = q[:, 0:1] * q[:, 1:4]
    U[:, 4] = q[:, 4] / (GAMMA - 1) + 0.5 * q[:, 0] * (q[:, 1:4] ** 2).sum(1) + (q[:, 5:8] ** 2).sum(1) / (8 * np.pi)
    U[:, 5:8] = q[:, 5:8]
    return U


def cell_errors(got, ref):
    """Per-cell relative errors of the conserved variables, each cell against itself: rho and E by
    |d|/|ref|, momentum and B by the norm of the vector difference over the cell's |rho v| or |B|;
    and the pressure by |dp|/p.  Returns the largest of each over the mesh."""
    Ug, Ur = conserved_state(got), conserved_state(ref)
    tiny = 1e-300
    e = dict(
        rho=np.abs(Ug[:, 0] - Ur[:, 0]) / np.maximum(np.abs(Ur[:, 0]), tiny),
        mom=np.linalg.norm(Ug[:, 1:4] - Ur[:, 1:4], axis=1) / np.maximum(np.linalg.norm(Ur[:, 1:4], axis=1), tiny),
        E=np.abs(Ug[:, 4] - Ur[:, 4]) / np.maximum(np.abs(Ur[:, 4]), tiny),
        B=np.linalg.norm(Ug[:, 5:8] - Ur[:, 5:8], axis=1) / np.maximum(np.linalg.norm(Ur[:, 5:8], axis=1), tiny),
        p=np.abs(np.asarray(got)[:, 4] - np.asarray(ref)[:, 4]) / np.maximum(np.abs(np.asarray(ref)[:, 4]), tiny))
    return {k: float(v.max()) for k, v in e.items()}


def cell_err(got, ref):
    """Largest per-cell relative error over the conserved variables (rho, rho v, E, B)."""
    e = cell_errors(got, ref)
    return max(e["rho"], e["mom"], e["E"], e["B"])


def node_err(got, ref):
    """Largest per-node relative position error |dx|/|x|."""
    return float((np.linalg.norm(np.asarray(got) - ref, axis=1) / np.maximum(np.linalg.norm(ref, axis=1), 1e-300)).max())


# ---- synthetic sets of the B, C, D, E, H, G rays (Luch.cpp:153-739) ------------------------------
EXT_M = dict(M0=3, M1=4, M11=2, M2=8, M3=5, M4=4, MF=2, N4=5, N5=6)


def ext_ray_need(t, g):
    """Nodes a ray of type t needs (the index count Luch::dvigenie runs over)."""
    M0, M1, M11, M2, M3, M4 = (g[k] for k in ("M0", "M1", "M11", "M2", "M3", "M4"))
    return {1: M0 + M1 + M2 + M3 + M4 + 5, 2: M0 + M1 + M11 + g["N4"] + g["N5"] + 2, 3: M2 - M11 + M3 + M4 + 4,
            4: M3 + M4 + 3, 5: 5, 6: M2 - M11 - g["MF"] + 2}[t]


def ext_ray_set(rng, first_node, n_nodes, counts=(6, 6, 6, 6, 6, 6), serial=(), geo=None):
    """Rays of each family over nodes first_node.. of a node array of n_nodes, laid out as the
    reference builds them (Setka.cpp:1600-1900): B rays run out along (the, phi) with support nodes
    at their own TS, HP and BS indices; C rays bend tailward from their TS node; D rays start at a
    node of a C ray's bend, E rays take their x from a C ray's tail node; H rays bridge a B ray's HP
    node to a D ray's, G rays a B ray's TS node to its HP node.  Families named in `serial` get rays
    that read a node the previous ray of the family writes.  Returns (types, node_lists, opor
    [n, 4], angles [n, 2], geo, xyz [n_nodes, 3] with plausible starting positions of the nodes
    used)."""
    g = dict(EXT_M, R0=1.0, R1=10.0, R5=600.0, L6=-60.0, L7=-160.0, dd1=1.0, dd2=1.0, dd3=19.0, dd4=15.0, dd5=19.0,
             dd6=14.0, dd7=0.7, dd8=1.0, ba1=1.3, ba2=0.8, ba3=0.7, ba4=1.6, ba5=0.3, md1=0.2, md2=0.5, md3=0.3,
             ea1=0.1, ea2=0.5, ea3=0.2)
    g.update(geo or {})
    M0, M1, M2, M3 = g["M0"], g["M1"], g["M2"], g["M3"]
    xyz = rng.uniform(-300.0, 300.0, (n_nodes, 3))
    nxt = [first_node]

    def take(n):
        a = nxt[0]
        nxt[0] += n
        assert nxt[0] <= n_nodes
        return list(range(a, a + n))

    def put(v, r, the, phi, jit=0.05):
        d = np.array([math.cos(the), math.sin(the) * math.cos(phi), math.sin(the) * math.sin(phi)])
        xyz[v] = r * d * (1.0 + jit * rng.uniform(-1, 1))

    def put_cyl(v, x, r, phi):
        xyz[v] = [x, r * math.cos(phi), r * math.sin(phi)]

    types, lists, opor, angles = [], [], [], []
    iTS, iHP, iBS = M0 + M1 + 1, M0 + M1 + M2 + 2, M0 + M1 + M2 + M3 + 3
    B, Cr, D = [], [], []
    for _ in range(counts[0]):
        the, phi = rng.uniform(1.2, 1.9), rng.uniform(0, 2 * math.pi)
        nodes = take(ext_ray_need(1, g))
        put(nodes[iTS], rng.uniform(75, 95), the, phi)
        put_cyl(nodes[iHP], rng.uniform(-20, 20), rng.uniform(110, 140), phi)
        put_cyl(nodes[iBS], rng.uniform(-20, 20), rng.uniform(220, 280), phi)
        types.append(1); lists.append(nodes); angles.append((the, phi))
        opor.append([nodes[M0], nodes[iTS], nodes[iHP], nodes[iBS]])
        B.append(nodes)
    c_bend = M0 + M1 + g["M11"] + 2
    for _ in range(counts[1]):
        the, phi = rng.uniform(2.0, 2.8), rng.uniform(0, 2 * math.pi)
        nodes = take(ext_ray_need(2, g))
        put(nodes[iTS], rng.uniform(85, 110), the, phi)
        types.append(2); lists.append(nodes); angles.append((the, phi))
        opor.append([nodes[M0], nodes[iTS], -1, -1])
        Cr.append(nodes)
    nD = ext_ray_need(3, g)
    for i in range(counts[2]):
        c = Cr[i % len(Cr)]
        phi = rng.uniform(0, 2 * math.pi)
        own = take(nD - 1)
        nodes = [c[c_bend + int(rng.integers(0, g["N4"]))]] + own
        hp, bs = nodes[M2 - g["M11"] + 1], nodes[M2 - g["M11"] + M3 + 2]
        put_cyl(hp, rng.uniform(-40, 0), rng.uniform(115, 145), phi)
        put_cyl(bs, rng.uniform(-40, 0), rng.uniform(230, 290), phi)
        types.append(3); lists.append(nodes); angles.append((0.0, phi))
        opor.append([nodes[0], hp, bs, -1])
        D.append(nodes)
    for i in range(counts[3]):
        c = Cr[i % len(Cr)]
        phi = rng.uniform(0, 2 * math.pi)
        nodes = take(ext_ray_need(4, g))
        put_cyl(nodes[0], rng.uniform(-120, -80), rng.uniform(120, 150), phi)
        put_cyl(nodes[M3 + 1], rng.uniform(-120, -80), rng.uniform(240, 300), phi)
        types.append(4); lists.append(nodes); angles.append((0.0, phi))
        opor.append([c[c_bend + g["N4"] + int(rng.integers(0, g["N5"]))], nodes[0], nodes[M3 + 1], -1])
    prev = None
    for i in range(counts[4]):
        b, d = B[i % len(B)], D[i % len(D)]
        own = take(4)
        z3 = b[iHP - 1] if not (5 in serial and prev is not None) else prev[2]
        z1, z2, z4 = b[iHP], d[M2 - g["M11"] + 1], d[M2 - g["M11"] + 2]
        types.append(5); lists.append([z1] + own + [z2]); angles.append((0.0, rng.uniform(0, 2 * math.pi)))
        opor.append([z3, z1, z4, z2])
        prev = own
    prev = None
    nG = g["M2"] - g["M11"] - g["MF"] + 1
    for i in range(counts[5]):
        b = B[i % len(B)]
        own = take(nG)
        z3 = b[iTS - 1] if not (6 in serial and prev is not None) else prev[1]
        types.append(6); lists.append([b[iTS]] + own + [b[iHP]]); angles.append((0.0, rng.uniform(0, 2 * math.pi)))
        opor.append([z3, b[iTS], b[iHP], -1])
        prev = own
    return (np.asarray(types, np.int8), lists, np.asarray(opor, np.int32), np.asarray(angles, np.float64), g,
            xyz)


# ---- HP grid lists on the fitted shell (Setka_move.cpp:700-1357 reads them off the ray grids) -----
def _wrap(a):
    return (a + math.pi) % (2 * math.pi) - math.pi


def hp_grid_lists(rays, xyz, band=0.12, the_lo=0.35, the_hi=2.8, n_grid=None, seed=3):
    """Ray-grid neighbour lists built from the shell's ray directions: for an HP node at polar angle
    the (from +x) and azimuth phi, its azimuthal neighbours are the HP nodes of the band |d the| <
    band ordered by azimuth (i-1, i-2, i+1, i+2, i+3, i-3 for the angle term; i-1, i-2, i+1, i+2 for
    macros1), its along-grid neighbours the HP nodes nearest in direction to (the -+ d, phi) and
    (the -+ 2d, phi).  Returns dict(angle_nodes, nb6, grid_nodes, nb8, tail [n_az, 5], bs_ref,
    bs_lists) in reference node ids; tails and BS lists are short latitude runs on the flank."""
    M0, M1, M2, M3 = rays["M0"], rays["M1"], rays["M2"], rays["M3"]
    iHP, iBS = M0 + M1 + M2 + 2, M0 + M1 + M2 + M3 + 3
    rn = np.asarray(rays["ray_nodes"])
    the, phi = rays["ray_angles"][:, 0], rays["ray_angles"][:, 1]
    hp, bs = rn[:, iHP], rn[:, iBS]
    unit = np.stack([np.cos(the), np.sin(the) * np.cos(phi), np.sin(the) * np.sin(phi)], axis=1)
    d = float(np.median(np.sort(np.arccos(np.clip(unit @ unit.T, -1, 1)), axis=1)[:, 1]))
    rng = np.random.default_rng(seed)
    ok = np.where((the > the_lo) & (the < the_hi))[0]
    angle_nodes, nb6, grid_nodes, nb8 = [], [], [], []
    for k in ok:
        ring = np.where(np.abs(the - the[k]) < band)[0]
        ring = ring[ring != k]
        if len(ring) < 8:
            continue
        dphi = _wrap(phi[ring] - phi[k])
        left = ring[dphi < 0][np.argsort(-dphi[dphi < 0])]
        right = ring[dphi > 0][np.argsort(dphi[dphi > 0])]
        if len(left) < 3 or len(right) < 3:
            continue
        angle_nodes.append(hp[k])
        nb6.append([hp[left[0]], hp[left[1]], hp[right[0]], hp[right[1]], hp[right[2]], hp[left[2]]])
        along = []
        for s in (-1, -2, 1, 2):
            t = the[k] + s * d
            tgt = np.array([math.cos(t), math.sin(t) * math.cos(phi[k]), math.sin(t) * math.sin(phi[k])])
            j = int(np.argmax(unit @ tgt))
            along.append(j)
        xs = [xyz[hp[j], 0] for j in along] + [xyz[hp[k], 0]]
        if len(set(along)) < 4 or k in along or len(set(np.round(xs, 12))) < 5:
            continue
        grid_nodes.append(hp[k])
        nb8.append([hp[along[0]], hp[along[1]], hp[along[2]], hp[along[3]], hp[left[0]], hp[left[1]], hp[right[0]],
                    hp[right[1]]])
    if n_grid is not None and len(grid_nodes) > n_grid:
        pick = np.sort(rng.choice(len(grid_nodes), n_grid, replace=False))
        grid_nodes = [grid_nodes[i] for i in pick]
        nb8 = [nb8[i] for i in pick]
    # tails: runs of 5 HP nodes across the flank (the from 80 to 100 degrees) at one azimuth;
    # BS lists: BS nodes of a latitude band, onto the first one's r_yz
    tail, bs_ref, bs_lists, used = [], [], [], set()
    for k in np.where(np.abs(the - 0.5 * math.pi) < 0.05)[0]:
        run = []
        for s in range(5):
            t = the[k] + (s - 2) * d
            tgt = np.array([math.cos(t), math.sin(t) * math.cos(phi[k]), math.sin(t) * math.sin(phi[k])])
            run.append(int(np.argmax(unit @ tgt)))
        if len(set(run)) == 5 and not (set(run) & used):
            used |= set(run)
            tail.append([hp[j] for j in run])
    for k in np.where(np.abs(the - 0.5 * math.pi) < 0.05)[0][:6]:
        band_k = np.where((np.abs(the - the[k]) < 0.5 * d) & (np.abs(_wrap(phi - phi[k])) < 3 * d))[0]
        band_k = [j for j in band_k if j != k and bs[j] not in bs_ref and all(bs[j] not in b for b in bs_lists)]
        if band_k and bs[k] not in [x for b in bs_lists for x in b]:
            bs_ref.append(bs[k])
            bs_lists.append([bs[j] for j in band_k[:3]])
    return dict(angle_nodes=np.asarray(angle_nodes, np.int32), nb6=np.asarray(nb6, np.int32),
                grid_nodes=np.asarray(grid_nodes, np.int32), nb8=np.asarray(nb8, np.int32),
                tail=np.asarray(tail[:8], np.int32).reshape(-1, 5), bs_ref=np.asarray(bs_ref, np.int32),
                bs_lists=bs_lists)
