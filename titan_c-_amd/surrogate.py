"""Charge-exchange surrogate: training-set generation, device training, installation.

The network maps 8 features of a cell
    x = [ln n_p, ln c_p^2, ln n_H, ln c_H^2, dV_x, dV_y, dV_z, |dV|],  dV = V_p - V_H,
to its 5 sources per unit density product,
    s = [rate, S_mx, S_my, S_mz, S_E'] / (n_p n_H),   S_E' = S_E - V_p . S_m (plasma frame),
in scaled residual form: output k is y_k = (s_k - b_k) / m_k, where b_k is a closed-form estimate
of s_k from the sigma = const moments of the relative-velocity Maxwellian (the baseline of
mlp_device.h, restated here in FP64) and m_k > 0 its magnitude.  The estimates are within a few
per cent, so the network learns small corrections and its BF16 arithmetic costs correspondingly
little accuracy.  Inputs are normalised by per-feature mean/std, outputs by a per-output scale.
Labels are the exact FP64 Maxwellian charge-exchange integrals computed on the GPU
(titan_cx_exact_batch); training runs on the GPU in the library's FP32 matrix-core kernels (forward,
backward and weight gradient fused per mini-batch, then one reduce + Adam kernel, mini-batches drawn
by a keyed permutation on the device); inference runs in the BF16 MFMA kernel with the exact source as the fallback for cells
outside the training box.
"""
import math

import numpy as np
from scipy.special import erf

from . import cx_exact_batch

N_FEAT, N_OUT = 8, 5


def sigma_ls(g):
    """The Lindsay-Stebbings cross-section of the exact source (cx_device.h), per code unit."""
    e_per_g2 = 0.5 * 1.67262192369e-24 * (1.03804e6 * 1.03804e6) / 1.602176634e-9
    nl16 = 1e-16 * 0.06 * 4.2113 * 1.495978707e13
    E = np.maximum(e_per_g2 * g * g, 1e-12)
    a = 4.15 - 0.531 * np.log(E)
    b = 1.0 - np.exp(-67.3 / E)
    return a * a * np.exp(4.5 * np.log(b)) * nl16


def baseline(U, c, kappa, e):
    """Closed-form estimates b [n, 5] of the sources per n_p n_H and their scales m [n, 5] (the
    formulas of mlp_baseline, mlp_device.h, which forms them in FP32 from hardware transcendentals
    and the erf of Abramowitz & Stegun 7.1.26; this FP64 restatement makes the labels): <g> and
    <g^3> of the relative-velocity Maxwellian of drift U and width c, U u_m with u_m^2 = U^2 +
    64 c^2 / (9 pi), sigma at <g>, u_m, <g^3>^(1/3)."""
    rpi = 1.0 / math.sqrt(math.pi)
    w = U / c
    small = w < 1e-3
    ws = np.where(small, 1.0, w)
    ew, erw = np.exp(-ws * ws) * rpi, erf(ws)
    g1 = np.where(small, c * 2 * rpi * (1 + w * w / 3), c * (ew + (ws + 0.5 / ws) * erw))
    g3 = np.where(small, c ** 3 * 4 * rpi * (1 + w * w),
                  c ** 3 * ((ws * ws + 2.5) * ew + (ws ** 3 + 3 * ws + 0.75 / ws) * erw))
    um = np.sqrt(U * U + 64.0 / (9.0 * math.pi) * c * c)
    B0 = sigma_ls(g1) * g1
    B1 = sigma_ls(um) * U * um
    B2 = sigma_ls(np.cbrt(g3)) * g3
    b = np.column_stack([B0, -B1[:, None] * e, 0.5 * (U * B1 * (1 + kappa) - kappa * B2)])
    m = np.column_stack([B0, B1, B1, B1, 0.5 * (U * B1 * (1 + kappa) + (np.abs(kappa) + 0.01) * B2)])
    return b, np.maximum(m, 1e-30)


def relative_frame(cp2, vp, ch2, vh):
    dv = vp - vh
    U = np.sqrt((dv[:, 0] * dv[:, 0] + dv[:, 1] * dv[:, 1]) + dv[:, 2] * dv[:, 2])
    c2 = cp2 + ch2
    e = np.where(U[:, None] > 0, dv / np.where(U > 0, U, 1.0)[:, None], np.array([1.0, 0.0, 0.0]))
    return U, np.sqrt(c2), (cp2 - ch2) / c2, e


def features(n_p, cp2, vp, n_h, ch2, vh):
    dv = vp - vh
    return np.column_stack([np.log(n_p), np.log(cp2), np.log(n_h), np.log(ch2), dv,
                            np.sqrt((dv[:, 0] * dv[:, 0] + dv[:, 1] * dv[:, 1]) + dv[:, 2] * dv[:, 2])])


def targets(src5, n_p, cp2, vp, n_h, ch2, vh):
    """The network's labels: (source per n_p n_H - closed-form estimate) / scale, the energy in the
    plasma frame."""
    nn = (n_p * n_h)[:, None]
    s = src5 / nn
    s[:, 4] = s[:, 4] - (vp * s[:, 1:4]).sum(axis=1)
    b, m = baseline(*relative_frame(cp2, vp, ch2, vh))
    return (s - b) / m


def sources(y, n_p, cp2, vp, n_h, ch2, vh):
    """Sources [n, 5] (rate, S_m, S_E in the lab frame) from unnormalised network outputs y."""
    b, m = baseline(*relative_frame(cp2, vp, ch2, vh))
    nn = (n_p * n_h)[:, None]
    s = nn * (b + m * y)
    s[:, 4] = s[:, 4] + (vp * s[:, 1:4]).sum(axis=1)
    return s


def sample_states(prim, neutrals, n, seed, jitter=0.25):
    """n training states: plasma states drawn from `prim` (cells of the run) and neutral moments
    drawn from `neutrals`, independently, each perturbed log-normally (densities, temperatures)
    and by +-jitter of the speed (velocities) so the training box covers the states a run
    visits.  Returns in10 [n, 10] = n_p, c_p^2, V_p, n_H, c_H^2, V_H."""
    rng = np.random.default_rng(seed)
    i = rng.integers(0, prim.shape[0], n)
    j = rng.integers(0, neutrals.shape[0], n)
    p = prim[i]
    h = neutrals[j]
    n_p = p[:, 0] * np.exp(rng.normal(0, jitter, n))
    cp2 = p[:, 4] / p[:, 0] * np.exp(rng.normal(0, jitter, n))
    vp = p[:, 1:4] * (1 + jitter * rng.uniform(-1, 1, (n, 1))) + jitter * rng.normal(0, 1, (n, 3))
    n_h = h[:, 0] * np.exp(rng.normal(0, jitter / 2, n))
    ch2 = h[:, 1] * np.exp(rng.normal(0, jitter / 2, n))
    vh = h[:, 2:5] * (1 + jitter * rng.uniform(-1, 1, (n, 1)))
    return np.column_stack([n_p, cp2, vp, n_h, ch2, vh])


def make_dataset(in10, nodes=100, chunk=250_000):
    out = np.empty((in10.shape[0], 5))
    for s in range(0, in10.shape[0], chunk):
        out[s:s + chunk] = cx_exact_batch(in10[s:s + chunk], nodes)
    x = features(in10[:, 0], in10[:, 1], in10[:, 2:5], in10[:, 5], in10[:, 6], in10[:, 7:10])
    y = targets(out, in10[:, 0], in10[:, 1], in10[:, 2:5], in10[:, 5], in10[:, 6], in10[:, 7:10])
    return x, y, out


class Normaliser:
    def __init__(self, x, y, box_quantile=0.0):
        self.mean = x.mean(axis=0)
        self.std = x.std(axis=0) + 1e-12
        self.scale = y.std(axis=0) + 1e-30
        lo_q, hi_q = box_quantile, 1.0 - box_quantile
        self.lo = np.quantile(x, lo_q, axis=0)
        self.hi = np.quantile(x, hi_q, axis=0)

    def nx(self, x):
        return ((x - self.mean) / self.std).astype(np.float32)

    def ny(self, y):
        return (y / self.scale).astype(np.float32)

    def blocks(self):
        in_norm = np.concatenate([self.mean, self.std]).astype(np.float32)
        box = np.concatenate([self.lo, self.hi]).astype(np.float32)
        return in_norm, self.scale.astype(np.float32), box


def init_params(seed):
    """He-style initialisation of W1..W4 (rows = outputs), zero biases."""
    rng = np.random.default_rng(seed)
    dims = [8, 64, 64, 64, 5]
    parts = []
    for a, b in zip(dims[:-1], dims[1:]):
        parts.append(rng.normal(0, np.sqrt(2.0 / a), (b, a)).reshape(-1))
        parts.append(np.zeros(b))
    return np.concatenate(parts).astype(np.float32)


def train(setka, x, y, norm, epochs=(8, 6, 4), lrs=(2e-3, 5e-4, 1e-4), batch=4096, seed=0):
    """Trains on the device; returns the per-epoch loss history."""
    in_norm, scale, box = norm.blocks()
    setka.mlp_set(init_params(seed), in_norm, scale, box)
    xn, yn = norm.nx(x), norm.ny(y)
    hist = []
    for k, (ep, lr) in enumerate(zip(epochs, lrs)):
        hist.extend(setka.mlp_train(xn, yn, ep, batch, lr, seed=seed + 17 * k).tolist())
    return np.array(hist)
