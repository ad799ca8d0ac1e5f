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
