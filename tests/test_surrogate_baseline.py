"""The closed-form estimates the surrogate network corrects (surrogate.baseline, the FP64
restatement of mlp_baseline in mlp_device.h).  Pinned by Monte Carlo: <g> and <g^3> of a shifted
Maxwellian are exact closed forms; U u_m = U sqrt(U^2 + 64 c^2 / (9 pi)) matches <g g_par> at both
ends (small and large drift) and within a few per cent between; and against the exact FP64
quadrature of the Lindsay-Stebbings source on states of the shell workload the estimates are
within a few per cent, so the network's labels are small corrections."""
import importlib
import math

import numpy as np


def mc_moments(U, c, n=2_000_000, seed=0):
    rng = np.random.default_rng(seed)
    R = rng.normal(0.0, c / math.sqrt(2.0), (n, 3))   # exp(-R^2/c^2)
    g = R.copy()
    g[:, 0] += U
    gm = np.linalg.norm(g, axis=1)
    return gm.mean(), (gm * g[:, 0]).mean(), (gm ** 3).mean()


def test_sigma_const_moments_against_monte_carlo():
    sur = importlib.import_module("titan_c-_amd.surrogate")
    one = lambda g: np.ones_like(g)
    orig = sur.sigma_ls
    sur.sigma_ls = one
    try:
        # small drift: <g g_par> -> U (4/3) <g>_0 = U 8c / (3 sqrt(pi)) (g ~ R + U cos, g_par = U + R cos,
        # averaged over directions), which U u_m reproduces
        U, c = 1e-4, 1.3
        b, m = sur.baseline(np.array([U]), np.array([c]), np.array([0.2]), np.array([[1.0, 0.0, 0.0]]))
        assert abs(-b[0, 1] / (U * 8 * c / (3 * math.sqrt(math.pi))) - 1) < 1e-7
        for U, c in ((0.3, 1.0), (1.0, 0.7), (3.0, 1.2), (12.0, 0.9)):
            b, m = sur.baseline(np.array([U]), np.array([c]), np.array([0.2]), np.array([[1.0, 0.0, 0.0]]))
            g1, g1par, g3 = mc_moments(U, c)
            # rate estimate = <g> exactly for sigma = 1
            assert abs(b[0, 0] / g1 - 1) < 2e-3, (U, c, b[0, 0], g1)
            # momentum: -U u_m against <g g_par> (an interpolation exact at both ends)
            assert abs(-b[0, 1] / g1par - 1) < 0.03, (U, c, -b[0, 1], g1par)
            # energy: B2 = <g^3> exactly, recovered from the plasma-frame estimate and its scale
            B1 = -b[0, 1]
            B2 = (0.5 * U * B1 * 1.2 - b[0, 4]) / (0.5 * 0.2)
            assert abs(B2 / g3 - 1) < 3e-3, (U, c, B2, g3)
    finally:
        sur.sigma_ls = orig


def test_baseline_within_a_few_percent_of_the_exact_source(C, O):
    sur = importlib.import_module("titan_c-_amd.surrogate")
    mesh, prim, params = C.shell(n_edge=6, n_layers=12)
    nh = C.neutrals(mesh.n_cells)
    in10 = sur.sample_states(prim, nh, 400, seed=4)
    exact = np.array([O.cx_source(a[0], a[1], a[2:5], a[5], a[6], a[7:10], nodes=100) for a in in10])
    y = sur.targets(exact, in10[:, 0], in10[:, 1], in10[:, 2:5], in10[:, 5], in10[:, 6], in10[:, 7:10])
    assert np.abs(y).max() < 0.08, np.abs(y).max(axis=0)
    # and the assembly inverts the labels
    back = sur.sources(y, in10[:, 0], in10[:, 1], in10[:, 2:5], in10[:, 5], in10[:, 6], in10[:, 7:10])
    assert np.abs(back - exact).max() <= 1e-12 * np.abs(exact).max()
