"""Pins the oracle's Phys_param::chlld (Phys_param.cpp:882-1878) by properties the mathematics
fixes.  The reference ships no golden vectors for it and cannot be built here (Eigen absent), so
each check is chosen so that a plausible restatement mistake (dropped term, wrong sign or index,
swapped region formula, wrong rotation) fails at least one of them:
  * consistency with the reference's own physical flux Get_Potok (Phys_param.cpp:826-874):
    chlld(q, q) = F(q).n - w U(q) for every method and every w;
  * rotational invariance and left/right mirror symmetry;
  * HLLD reduces to HLLC when B.n = 0;
  * the region index is monotone in the face speed w and the ALE flux F - wU is continuous
    across every wave (mass, momentum, energy);
  * get_bazis returns an orthonormal right-handed frame.
"""
import numpy as np
import pytest

from helpers import random_normals, random_rotation, random_states

SPI4 = np.sqrt(4.0 * np.pi)
GAMMA = 5.0 / 3.0


def conserved(q):
    """U = (rho, rho v, E, B) in the chlld output order (energy at index 4)."""
    rho, v, p, B = q[0], q[1:4], q[4], q[5:8]
    E = p / (GAMMA - 1) + 0.5 * rho * v @ v + (B @ B) / (8 * np.pi)
    return np.array([rho, *(rho * v), E, *B])


def test_get_bazis_orthonormal(O):
    rng = np.random.default_rng(0)
    for n in list(random_normals(rng, 200)) + [np.array([1.0, 0, 0]), np.array([0, 0, 1.0]),
                                                 np.array([-0.95, 0.1, np.sqrt(1 - 0.9125)])]:
        t, m = O.get_bazis(n)
        M = np.stack([n, t, m])
        assert np.abs(M @ M.T - np.eye(3)).max() < 1e-14
        assert abs(np.linalg.det(M) - 1.0) < 1e-14


@pytest.mark.parametrize("method", [0, 1, 2, 3])
def test_consistency_with_get_potok(O, method):
    rng = np.random.default_rng(1 + method)
    n = 400
    q = random_states(rng, n)
    nrm = random_normals(rng, n)
    w = rng.normal(0, 3.0, n)
    w[:100] = 0.0
    got, _, _, br, rc = O.chlld_batch(method, nrm, q, q, w=w)
    assert rc == 0
    for i in range(n):
        pot = O.get_potok(GAMMA, q[i], nrm[i])
        F = np.array([pot[0], pot[1], pot[2], pot[3], pot[7], pot[4], pot[5], pot[6]])
        U = conserved(q[i])
        exp = F - w[i] * U
        scale = np.abs(F).max() + abs(w[i]) * np.abs(U).max()
        assert np.abs(got[i] - exp).max() <= 1e-12 * scale, (method, i, got[i], exp)


@pytest.mark.parametrize("method", [1, 2, 3])
def test_rotational_invariance(O, method):
    rng = np.random.default_rng(10 + method)
    n = 300
    qL, qR = random_states(rng, n), random_states(rng, n)
    nrm = random_normals(rng, n)
    w = rng.normal(0, 1.0, n)
    f0, _, ds0, br0, _ = O.chlld_batch(method, nrm, qL, qR, w=w)
    R = random_rotation(rng)

    def rot(q):
        q2 = q.copy()
        q2[:, 1:4] = q[:, 1:4] @ R.T
        q2[:, 5:8] = q[:, 5:8] @ R.T
        return q2

    f1, _, ds1, br1, _ = O.chlld_batch(method, nrm @ R.T, rot(qL), rot(qR), w=w)
    scale = np.abs(f0).max(axis=1)
    assert (np.abs(rot(f0) - f1).max(axis=1) / scale).max() < 1e-11
    assert np.abs(ds0 - ds1).max() < 1e-11 * np.abs(ds0).max()


@pytest.mark.parametrize("method", [0, 1, 2, 3])
def test_mirror_symmetry(O, method):
    """F(qL, qR; n, w) = -F(qR, qL; -n, -w)."""
    rng = np.random.default_rng(20 + method)
    n = 300
    qL, qR = random_states(rng, n), random_states(rng, n)
    nrm = random_normals(rng, n)
    w = rng.normal(0, 1.0, n)
    f0, _, ds0, b0, _ = O.chlld_batch(method, nrm, qL, qR, w=w)
    f1, _, ds1, b1, _ = O.chlld_batch(method, -nrm, qR, qL, w=-w)
    scale = np.abs(f0).max(axis=1)
    # The reference's HLLD tests its regions with separate ifs (Phys_param.cpp:1688-1832): when
    # an Alfven-speed estimate S*_L (S*_R) falls outside [S_L, S_M] ([S_M, S_R]) a double-star
    # region can overwrite the outer one for w outside the fan, which breaks the symmetry.  Those
    # faces are excluded here and pinned by test_hlld_overwrite_outside_fan.
    outside = (w < ds0[:, 0]) | (w > ds0[:, 2])
    anomal = outside & (np.isin(b0 & 7, [3, 4]) | np.isin(b1 & 7, [3, 4])) if method == 3 else np.zeros(n, bool)
    assert anomal.mean() < 0.02
    ok = ~anomal
    assert (np.abs(f0 + f1).max(axis=1)[ok] / scale[ok]).max() < 1e-11


def test_hlld_reduces_to_hllc_without_normal_field(O):
    rng = np.random.default_rng(30)
    n = 300
    qL, qR = random_states(rng, n), random_states(rng, n)
    nrm = np.tile([1.0, 0.0, 0.0], (n, 1))
    qL[:, 5] = 0.0
    qR[:, 5] = 0.0
    w = rng.normal(0, 1.0, n)
    fd, _, _, brd, _ = O.chlld_batch(3, nrm, qL, qR, w=w)
    fc, _, _, brc, _ = O.chlld_batch(2, nrm, qL, qR, w=w)
    assert np.all((brd >> 3) & 1 == 0)   # ibn = 0
    scale = np.abs(fc).max(axis=1)
    assert (np.abs(fd - fc).max(axis=1) / scale).max() < 1e-12


@pytest.mark.parametrize("method", [1, 2, 3])
def test_regions_monotone_and_flux_continuous(O, method):
    rng = np.random.default_rng(40 + method)
    for trial in range(40):
        qL, qR = random_states(rng, 1), random_states(rng, 1)
        nrm = random_normals(rng, 1)
        # a physical face has a continuous normal field (the momentum flux of the star states
        # jumps by B_nR^2 - B_nL^2 across the contact otherwise)
        qR[0, 5:8] += (qL[0, 5:8] @ nrm[0] - qR[0, 5:8] @ nrm[0]) * nrm[0]
        _, _, ds, _, _ = O.chlld_batch(method, nrm, qL, qR)
        lo, hi = ds[0, 0] - 1.0, ds[0, 2] + 1.0
        ws = np.linspace(lo, hi, 4001)
        m = ws.size
        f, _, _, br, rc = O.chlld_batch(method, np.repeat(nrm, m, 0), np.repeat(qL, m, 0), np.repeat(qR, m, 0), w=ws)
        assert rc == 0
        reg = br & 7
        assert np.all(np.diff(reg.astype(int)) >= 0), reg
        assert reg[0] == 1 and reg[-1] == (3 if method == 1 else (4 if method == 2 else 6))
        # continuity of every flux component across each region change: the ALE flux F - wU of
        # each side's HLL-type state is continuous at its outer waves by construction, and at the
        # contact only when the star and double-star states satisfy the jump conditions there, so
        # the tangential-B components (6, 7) pin the double-star field
        assert fan_discontinuity(f, reg) == 0, (method, trial)


def fan_discontinuity(f, reg):
    """Number of region changes across which some flux component jumps by more than the local
    slope allows."""
    jumps = np.nonzero(np.diff(reg.astype(int)))[0]
    bad = 0
    for j in jumps:
        for comp in (slice(0, 5), slice(5, 8)):
            scale = np.abs(f[:, comp]).max()
            d = np.abs(f[j + 1, comp] - f[j, comp]).max()
            slope = max(np.abs(f[j, comp] - f[j - 1, comp]).max(), np.abs(f[j + 2, comp] - f[j + 1, comp]).max())
            bad += int(d > 4 * slope + 1e-9 * scale)
    return bad


def test_fatal_branch_on_nan(O):
    q = np.array([[1.0, 0, 0, 0, np.nan, 0, 0, 0]])
    for method in range(4):
        _, _, _, br, rc = O.chlld_batch(method, np.array([[1.0, 0, 0]]), q, q)
        assert rc == -1
        assert br[0] & 7 == 0


def test_raspad_testing_inputs_are_admissible(O):
    """Inputs of the reference's manual smoke test Phys_param::raspad_testing
    (Phys_param.cpp:1892-1926): L=(1,1,1,1,1,1.1,1.2,1.3), R=(2,-1,1,1,6,3.1,1.2,1.3), n=x, w=-4,
    HLLC.  The reference prints the flux without expected values; the oracle must pass its
    consistency checks on them (w = -4 lies left of the fan: region 1, F(qL) - w U(qL))."""
    qL = np.array([[1, 1, 1, 1, 1, 1.1, 1.2, 1.3]], float)
    qR = np.array([[2, -1, 1, 1, 6, 3.1, 1.2, 1.3]], float)
    for w in (-4.0, -40.0):
        f, _, ds, br, rc = O.chlld_batch(2, np.array([[1.0, 0, 0]]), qL, qR, w=np.array([w]))
        assert rc == 0 and np.all(np.isfinite(f))
        if w < ds[0, 0]:
            # left of the whole fan: region 1, F(qL) - w U(qL) except the normal-B component,
            # which carries the reference's own Bn-flux formula
            assert br[0] & 7 == 1
            pot = O.get_potok(GAMMA, qL[0], np.array([1.0, 0, 0]))
            F = np.array([pot[0], pot[1], pot[2], pot[3], pot[7], pot[4], pot[5], pot[6]])
            exp = F - w * conserved(qL[0])
            mask = np.array([1, 1, 1, 1, 1, 0, 1, 1], bool)
            assert np.abs(f[0][mask] - exp[mask]).max() < 1e-12 * np.abs(exp).max()
        else:
            assert 1 <= br[0] & 7 <= 4


def test_hlld_overwrite_outside_fan(O):
    """Reference behaviour kept on purpose: for this face the mirrored solve has w < S_L, so region
    1 matches first, but S*_L < w <= S_M also holds and region 3 overwrites it."""
    rng = np.random.default_rng(23)
    n = 300
    qL, qR = random_states(rng, n), random_states(rng, n)
    nrm = random_normals(rng, n)
    w = rng.normal(0, 1.0, n)
    i = 250
    _, _, ds, br, _ = O.chlld_batch(3, -nrm[i:i + 1], qR[i:i + 1], qL[i:i + 1], w=-w[i:i + 1])
    assert -w[i] < ds[0, 0]          # left of the fan ...
    assert br[0] & 7 == 3            # ... yet the double-star region was selected


# ---- HLLD on isolated discontinuities (Miyoshi & Kusano, J. Comput. Phys. 208 (2005) 315) --------
# HLLD resolves an isolated contact and an isolated rotational discontinuity exactly: the ALE flux
# is F(U_L) - w U_L on the left of the discontinuity's speed and F(U_R) - w U_R on its right.  F is
# the textbook ideal-MHD flux in Gaussian units, written here independently of the oracle.

def mhd_flux(q, n):
    rho, v, p, B = q[0], q[1:4], q[4], q[5:8]
    vn, Bn = v @ n, B @ n
    pt = p + B @ B / (8 * np.pi)
    E = p / (GAMMA - 1) + 0.5 * rho * v @ v + B @ B / (8 * np.pi)
    mom = rho * v * vn + pt * n - Bn * B / (4 * np.pi)
    en = (E + pt) * vn - Bn * (v @ B) / (4 * np.pi)
    ind = vn * B - Bn * v
    return np.array([rho * vn, *mom, en, *ind])


def frame(n):
    a = np.array([1.0, 0.0, 0.0]) if abs(n[0]) < 0.9 else np.array([0.0, 1.0, 0.0])
    t = np.cross(a, n)
    t /= np.linalg.norm(t)
    return t, np.cross(n, t)


def isolated_error(O, qL, qR, nrm, speed, ws, method=3):
    """Largest error of chlld's flux against the exact flux of a single discontinuity moving at
    `speed`, over face speeds ws, relative to the flux scale."""
    m = ws.size
    f, _, _, br, rc = O.chlld_batch(method, np.tile(nrm, (m, 1)), np.tile(qL, (m, 1)), np.tile(qR, (m, 1)), w=ws)
    assert rc == 0
    err = 0.0
    for i, w in enumerate(ws):
        q = qL if w < speed else qR
        exact = mhd_flux(q, nrm) - w * conserved(q)
        scale = np.abs(mhd_flux(q, nrm)).max() + abs(w) * np.abs(conserved(q)).max()
        err = max(err, np.abs(f[i] - exact).max() / scale)
    return err, br


def contact_case(rng):
    nrm = random_normals(rng, 1)[0]
    t, m = frame(nrm)
    qL = random_states(rng, 1)[0]
    qL[5:8] += (0.5 + abs(qL[5:8] @ nrm)) * SPI4 * nrm       # |B_n| well above epsb
    qR = qL.copy()
    qR[0] = qL[0] * np.exp(rng.uniform(-2, 2))                 # only the density jumps
    return qL, qR, nrm, qL[1:4] @ nrm


def rotational_case(rng, right_going):
    nrm = random_normals(rng, 1)[0]
    t, m = frame(nrm)
    q = random_states(rng, 1)[0]
    rho = q[0]
    Bn = (0.5 + rng.uniform(0, 2)) * SPI4 * rng.choice([-1.0, 1.0])
    Bt = rng.normal(0, 2.0, 2) * SPI4
    phi = rng.uniform(0.3, 2.8)
    c, s = np.cos(phi), np.sin(phi)
    BtR = np.array([c * Bt[0] - s * Bt[1], s * Bt[0] + c * Bt[1]])   # rotated, same magnitude
    vt = rng.normal(0, 1.0, 2)
    sg = np.sign(Bn) * (1.0 if right_going else -1.0)
    vtR = vt - sg * (BtR - Bt) / np.sqrt(4 * np.pi * rho)             # Rankine-Hugoniot across it
    vn = rng.normal(0, 1.0)
    qL, qR = q.copy(), q.copy()
    qL[1:4] = vn * nrm + vt[0] * t + vt[1] * m
    qR[1:4] = vn * nrm + vtR[0] * t + vtR[1] * m
    qL[5:8] = Bn * nrm + Bt[0] * t + Bt[1] * m
    qR[5:8] = Bn * nrm + BtR[0] * t + BtR[1] * m
    speed = vn + (1.0 if right_going else -1.0) * abs(Bn) / np.sqrt(4 * np.pi * rho)
    return qL, qR, nrm, speed


def hlld_isolated_errors(O, seed=60, trials=30):
    rng = np.random.default_rng(seed)
    e_c = e_r = 0.0
    for _ in range(trials):
        qL, qR, nrm, u = contact_case(rng)
        ws = u + np.concatenate([-np.geomspace(5, 1e-3, 6), np.geomspace(1e-3, 5, 6)])
        e, br = isolated_error(O, qL, qR, nrm, u, ws)
        assert np.all((br >> 3) & 1 == 1)      # ibn = 1: the Alfven branch of HLLD is taken
        e_c = max(e_c, e)
        for rg in (True, False):
            qL, qR, nrm, lam = rotational_case(rng, rg)
            ws = lam + np.concatenate([-np.geomspace(5, 1e-3, 6), np.geomspace(1e-3, 5, 6)])
            e, br = isolated_error(O, qL, qR, nrm, lam, ws)
            e_r = max(e_r, e)
    return e_c, e_r


def test_hlld_isolated_contact_and_rotational_discontinuity(O):
    e_c, e_r = hlld_isolated_errors(O)
    assert e_c < 1e-12 and e_r < 1e-12, (e_c, e_r)


def test_hll_does_not_resolve_the_contact(O):
    """The contact check discriminates: HLL smears the contact (mass flux at the contact's speed
    of order S_L S_R (rho_R - rho_L)/(S_R - S_L)), HLLD gives the exact zero."""
    rng = np.random.default_rng(61)
    qL, qR, nrm, u = contact_case(rng)
    e_hll, _ = isolated_error(O, qL, qR, nrm, u, np.array([u + 1e-3]), method=1)
    e_hlld, _ = isolated_error(O, qL, qR, nrm, u, np.array([u + 1e-3]), method=3)
    assert e_hll > 1e-3 and e_hlld < 1e-12, (e_hll, e_hlld)


def test_hlld_pins_catch_planted_mistakes(O):
    """Planted errors of the star and double-star tangential states (oracle's test-only switch):
    the sign(Bn) term of v_t**, the sqrt(rho*) weights of b_t** and the Bn^2 of b_t* each fail the
    isolated-discontinuity check or the tangential-B flux continuity at the contact."""
    with O.planted_mistake(O.MISTAKE_HLLD_VZZ_SIGN):
        _, e_r = hlld_isolated_errors(O, trials=5)
    assert e_r > 1e-3, e_r
    with O.planted_mistake(O.MISTAKE_HLLD_TB_BN2):
        e_c, e_r = hlld_isolated_errors(O, trials=5)
    assert max(e_c, e_r) > 1e-3, (e_c, e_r)
    rng = np.random.default_rng(43)
    bad = 0
    with O.planted_mistake(O.MISTAKE_HLLD_BZZ_WEIGHTS):
        for trial in range(10):
            qL, qR = random_states(rng, 1), random_states(rng, 1)
            nrm = random_normals(rng, 1)
            qR[0, 5:8] += (qL[0, 5:8] @ nrm[0] - qR[0, 5:8] @ nrm[0]) * nrm[0]
            _, _, ds, _, _ = O.chlld_batch(3, nrm, qL, qR)
            ws = np.linspace(ds[0, 0] - 1.0, ds[0, 2] + 1.0, 4001)
            m = ws.size
            f, _, _, br, _ = O.chlld_batch(3, np.repeat(nrm, m, 0), np.repeat(qL, m, 0), np.repeat(qR, m, 0), w=ws)
            bad += fan_discontinuity(f, br & 7) > 0
    assert bad >= 8, bad
