"""Charge-exchange surrogate: training-set generation, device training, installation.

The network maps 8 features of a cell
    x = [ln n_p, ln c_p^2, ln n_H, ln c_H^2, dV_x, dV_y, dV_z, |dV|],  dV = V_p - V_H,
to 5 outputs per unit density product,
    y = [rate, S_mx, S_my, S_mz, S_E'] / (n_p n_H),   S_E' = S_E - V_p . S_m (plasma frame),
normalised by per-feature mean/std (inputs) and per-output scale.  Labels are the exact FP64
Maxwellian charge-exchange integrals computed on the GPU (titan_cx_exact_batch); training runs on
the GPU through the library's FP32 forward/backward/Adam kernels; inference runs in the BF16 MFMA
kernel with the exact source as the fallback for cells outside the training box.
"""
import numpy as np

from . import cx_exact_batch

N_FEAT, N_OUT = 8, 5


def features(n_p, cp2, vp, n_h, ch2, vh):
    dv = vp - vh
    return np.column_stack([np.log(n_p), np.log(cp2), np.log(n_h), np.log(ch2), dv,
                            np.sqrt((dv[:, 0] * dv[:, 0] + dv[:, 1] * dv[:, 1]) + dv[:, 2] * dv[:, 2])])


def targets(src5, n_p, n_h, vp):
    nn = (n_p * n_h)[:, None]
    y = src5 / nn
    y[:, 4] = y[:, 4] - (vp * y[:, 1:4]).sum(axis=1)
    return y


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
    y = targets(out, in10[:, 0], in10[:, 5], in10[:, 2:5])
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
