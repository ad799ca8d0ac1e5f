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
        # continuity of mass, momentum and energy fluxes across each region change
        jumps = np.nonzero(np.diff(reg.astype(int)))[0]
        h = ws[1] - ws[0]
        scale = np.abs(f[:, :5]).max()
        for j in jumps:
            d = np.abs(f[j + 1, :5] - f[j, :5]).max()
            slope = max(np.abs(f[j, :5] - f[j - 1, :5]).max(), np.abs(f[j + 2, :5] - f[j + 1, :5]).max())
            assert d <= 4 * slope + 1e-9 * scale, (method, trial, j, reg[j], reg[j + 1], d, slope)


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
