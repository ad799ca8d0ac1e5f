"""Parity of the device Riemann solvers (titan_chlld_batch) with the oracle's Phys_param::chlld."""
import numpy as np
import pytest

from helpers import random_normals, random_states

pytestmark = pytest.mark.gpu


def _case(rng, n):
    qL = random_states(rng, n)
    qR = random_states(rng, n)
    # a third of the faces get near-equal states (smooth flow), some get |Bn| below epsb
    k = n // 3
    qR[:k] = qL[:k] * (1.0 + 1e-3 * rng.normal(size=(k, 8)))
    nrm = random_normals(rng, n)
    w = rng.normal(0.0, 2.0, n)
    w[: n // 4] = 0.0
    return qL, qR, nrm, w


@pytest.mark.parametrize("method", [0, 1, 2, 3])
def test_chlld_fp64_bit_exact(T, O, method):
    rng = np.random.default_rng(100 + method)
    n = 250_000
    qL, qR, nrm, w = _case(rng, n)
    n_disc = rng.integers(0, 4, n).astype(np.int8)
    flags = rng.choice(np.array([0, 0, 0, 1, 2], np.uint8), n)
    ref = O.chlld_batch(method, nrm, qL, qR, w=w, n_disc=n_disc, flags=flags)
    got = T.chlld(method, nrm, qL, qR, w=w, n_disc=n_disc, flags=flags)
    ok = (ref[3] & 7) != 0  # region 0 = fatal branch, excluded from the flux comparison
    np.testing.assert_array_equal(got[3], ref[3])          # branch codes: n_state, ibn, region
    np.testing.assert_array_equal(got[2], ref[2])          # dsl, dsc, dsr
    np.testing.assert_array_equal(got[0][ok], ref[0][ok])  # fluxes, bit for bit


def test_chlld_passive_scalars_bit_exact(T, O):
    rng = np.random.default_rng(7)
    n = 60_000
    qL, qR, nrm, w = _case(rng, n)
    kL = np.abs(rng.normal(1.0, 0.5, (n, 2)))
    kR = np.abs(rng.normal(1.0, 0.5, (n, 2)))
    for method in range(4):
        ref = O.chlld_batch(method, nrm, qL, qR, w=w, kL=kL, kR=kR)
        got = T.chlld(method, nrm, qL, qR, w=w, kL=kL, kR=kR)
        ok = (ref[3] & 7) != 0
        np.testing.assert_array_equal(got[1][ok], ref[1][ok])
        np.testing.assert_array_equal(got[0][ok], ref[0][ok])


def test_chlld_brio_wu_interface(T, O):
    s4 = np.sqrt(4 * np.pi)
    qL = np.array([[1.0, 0, 0, 0, 1.0, 0.75 * s4, 1.0 * s4, 0.0]])
    qR = np.array([[0.125, 0, 0, 0, 0.1, 0.75 * s4, -1.0 * s4, 0.0]])
    for method in range(4):
        for w in (0.0, -0.3, 0.7):
            ref = O.chlld_batch(method, np.array([[1.0, 0, 0]]), qL, qR, w=np.array([w]))
            got = T.chlld(method, np.array([[1.0, 0, 0]]), qL, qR, w=np.array([w]))
            np.testing.assert_array_equal(got[0], ref[0])
            np.testing.assert_array_equal(got[3], ref[3])


def test_chlld_fp32_close(T, O):
    rng = np.random.default_rng(11)
    n = 100_000
    qL, qR, nrm, w = _case(rng, n)
    for method in (2, 3):
        ref = O.chlld_batch(method, nrm, qL, qR, w=w)
        got = T.chlld(method, nrm, qL, qR, w=w, precision=T.PREC_MIXED)
        scale = np.abs(ref[0]).max(axis=1, keepdims=True) + 1e-30
        err = np.abs(got[0] - ref[0]) / scale
        # FP32 Riemann arithmetic: errors relative to the face's largest flux component
        assert np.quantile(err, 0.999) < 1e-4, np.quantile(err, 0.999)
        assert (got[3] == ref[3]).mean() > 0.999


@pytest.mark.parametrize("method", [0, 1, 2, 3])
def test_chlld_mixed_error_scales_with_the_jump(T, O, method):
    """The mixed solver keeps the upwind flux in FP64 and forms every wave correction from the
    jumps: for states 1e-3 apart its error is ~1e-10 of the flux (a plain FP32 solve is ~1e-7),
    and with both states supersonic through the face it equals the FP64 flux to rounding."""
    rng = np.random.default_rng(200 + method)
    n = 60_000
    qL = random_states(rng, n)
    qR = qL * (1.0 + 1e-3 * rng.normal(size=(n, 8)))
    nrm = random_normals(rng, n)
    w = rng.normal(0.0, 0.5, n)
    ref = O.chlld_batch(method, nrm, qL, qR, w=w)
    got = T.chlld(method, nrm, qL, qR, w=w, precision=T.PREC_MIXED)
    scale = np.abs(ref[0]).max(axis=1, keepdims=True) + 1e-30
    err = (np.abs(got[0] - ref[0]) / scale).max(axis=1)
    assert np.quantile(err, 0.999) < 1e-8, np.quantile(err, 0.999)
    if method == 0:
        return  # Lax's speeds are symmetric about zero: it has no upwind region
    # supersonic: push both states through the face at 20x their speed scale along the normal
    vfast = 20.0 * (np.abs(qL[:, 1:4]).max(axis=1) + np.sqrt(qL[:, 4] / qL[:, 0]) + np.abs(qL[:, 5:8]).max(axis=1))
    qLs, qRs = qL.copy(), qR.copy()
    qLs[:, 1:4] += vfast[:, None] * nrm
    qRs[:, 1:4] += vfast[:, None] * nrm
    ref = O.chlld_batch(method, nrm, qLs, qRs, w=np.zeros(n))
    got = T.chlld(method, nrm, qLs, qRs, w=np.zeros(n), precision=T.PREC_MIXED)
    up = (ref[3] & 7) == 1  # the region left of the fan: the reference returns F_L
    assert up.mean() > 0.99
    scale = np.abs(ref[0][up]).max(axis=1, keepdims=True)
    # FP64 except the normal-field dissipation -S_N dBn, whose fast speed is FP32
    assert (np.abs(got[0][up][:, :5] - ref[0][up][:, :5]) / scale).max() < 1e-14
    assert (np.abs(got[0][up] - ref[0][up]) / scale).max() < 1e-12


@pytest.mark.parametrize("method", [0, 1, 2, 3])
def test_chlld_mixed_field_along_the_normal_at_the_alfven_speed(T, O, method):
    """Field along the normal and the sound speed a hair above the Alfven speed: the fast speed's
    inner radicand b^2/rho + c (c - 2 b_n/sqrt(rho)) = (c - b_n/sqrt(rho))^2 + (b^2 - b_n^2)/rho is
    ~1e-11 of c^2, which FP64 keeps and FP32 rounding takes below zero about half the time.  Every
    face still solves in the mixed solver (no fatal branch) and stays close to the FP64 flux."""
    rng = np.random.default_rng(31 + method)
    n = 20_000
    s4 = np.sqrt(4.0 * np.pi)
    gamma = 5.0 / 3.0
    rho = rng.uniform(0.01, 2.0, n)
    bn = rng.uniform(0.2, 3.0, n) * rng.choice([-1.0, 1.0], n)  # b = B / sqrt(4 pi)
    p = bn ** 2 / gamma * (1.0 + 1e-5 * rng.uniform(0.1, 1.0, n))  # c^2 = (b_n^2 / rho)(1 + eps)
    nrm = random_normals(rng, n)
    qL = np.zeros((n, 8))
    qL[:, 0], qL[:, 4] = rho, p
    qL[:, 5:8] = (bn * s4)[:, None] * nrm
    qL[:, 1:4] = rng.normal(0.0, 3.0, (n, 3))
    qR = qL.copy()
    qR[:, 1:4] += rng.normal(0.0, 0.05, (n, 3))  # a velocity jump; the radicands stay degenerate
    # the FP32 radicand on the host, in the solver's operation order, for the side whose sign of
    # b_n makes it cancel: a good share of the faces have it below zero
    f = np.float32
    aa = (bn * np.sign(bn)).astype(f) / np.sqrt(rho.astype(f))
    c = np.sqrt(f(gamma) * p.astype(f) / rho.astype(f))
    rad = (bn ** 2).astype(f) / rho.astype(f) + c * (c - f(2.0) * aa)
    assert (rad < 0).mean() > 0.1
    ref = O.chlld_batch(method, nrm, qL, qR, w=np.zeros(n))
    assert (ref[3] & 7).min() > 0  # FP64 solves every face
    got = T.chlld(method, nrm, qL, qR, w=np.zeros(n), precision=T.PREC_MIXED)
    assert got[4] == T.TITAN_OK
    assert (got[3] & 7).min() > 0
    scale = np.abs(ref[0]).max(axis=1, keepdims=True)
    assert (np.abs(got[0] - ref[0]) / scale).max() < 1e-4
