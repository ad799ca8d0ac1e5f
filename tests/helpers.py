"""Shared generators for the test suite."""
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
    """Error per conserved variable, relative to that variable's largest magnitude on the mesh."""
    return (np.abs(a - b).max(axis=0) / np.maximum(np.abs(b).max(axis=0), 1e-300)).max()
