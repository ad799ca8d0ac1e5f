"""GPU parity of the workload shapes and branches the other GPU tests do not reach: the
precision form on a prescribed-motion shell for 100 steps (the 4M-cell precision workload's
arithmetic at test size), the fused surrogate source on a moving shell, the update's density and
pressure floors (Setka_physics.cpp:1784-1794,1981-2001), and a titan_go call that fails its input
checks without touching the context.  Errors are per cell: each conserved variable against the
cell's own value."""
import importlib

import numpy as np
import pytest

from helpers import cell_err, cell_errors, node_err

pytestmark = pytest.mark.gpu


def moving_shell(C, n_edge=6, n_layers=10):
    return C.shell(n_edge=n_edge, n_layers=n_layers, moving=True, r_zone1_au=80.0, noise=0.01)


def test_mixed_precision_moving_shell_100_steps(T, C, O):
    """FP32 Riemann arithmetic with FP64 states, accumulation, pressure recovery and moving-mesh
    geometry, 100 prescribed-motion steps, against the FP64 oracle: per-cell relative error of
    every conserved variable and of the pressure."""
    mesh, prim, params = moving_shell(C)
    with T.Setka(mesh, dict(params, precision=T.PREC_MIXED)) as s:
        s.set_state(prim, 0)
        st = s.Go(100)
        got = s.get_state()
        xyz = s.get_nodes(s.current_layer())
    rc, ref, rxyz, dth, _ = O.go(O.OracleMesh(mesh), params, mesh.node_xyz, prim, 100, nthreads=8,
                                 disp=mesh.node_disp)
    assert rc == 0
    e = cell_errors(got, ref)
    print("mixed vs FP64, moving shell, 100 steps: per-cell errors", e, "nodes", node_err(xyz, rxyz),
          "dt", st["dt_last"], dth[-1], "FP64 re-solved faces (last step)", st["n_fp64_faces"])
    # the lagged dt follows the mixed fluxes' wave speeds (1e-8 relative here), and the nodes move
    # by dt times the motion law
    assert abs(st["dt_last"] / dth[-1] - 1) < 1e-6
    assert node_err(xyz, rxyz) < 1e-8
    # measured: conserved variables 6e-8, pressure 2.2e-5 (recovered from the total energy, which
    # the kinetic energy dominates in the cold supersonic wind)
    assert max(e["rho"], e["mom"], e["E"], e["B"]) < 1e-6, e
    assert e["p"] < 1e-4, e


@pytest.mark.parametrize("precision", [0, 1])
def test_fused_surrogate_moving_shell(T, C, O, precision):
    """One step of the fused surrogate source on the prescribed-motion shell: the moving-mesh
    branch of the surrogate update (old and new volumes, face velocities) equals the oracle's update
    of the moved geometry fed the standalone surrogate kernel's sources; cells whose ln n_p is above
    the median leave the training box and take the exact source.  FP64: 1e-12 per cell; FP32
    Riemann arithmetic: 1e-6."""
    sur = importlib.import_module("titan_c-_amd.surrogate")
    mesh, prim, params = moving_shell(C)
    nh = C.neutrals(mesh.n_cells)
    x = sur.features(prim[:, 0], prim[:, 4] / prim[:, 0], prim[:, 1:4], nh[:, 0], nh[:, 1], nh[:, 2:5])
    norm = sur.Normaliser(x, np.ones((len(x), 5)))
    in_norm, _, box = norm.blocks()
    box = box.copy()
    box[8] = np.float32(np.median(x[:, 0]))
    scale = np.full(5, 0.1, np.float32)
    P = sur.init_params(5)
    with T.Setka(mesh, dict(params, source_mode=T.SRC_MLP, cx_nodes=32, precision=precision)) as s:
        s.set_state(prim, 0)
        s.set_neutrals(nh)
        s.mlp_set(P, in_norm, scale, box)
        src = s.sources(0, T.SRC_MLP)
        st = s.Go(1)
        got = s.get_state()
        xyz = s.get_nodes(s.current_layer())
    assert st["n_fallback"] > 0
    om = O.OracleMesh(mesh)
    dt = 1e-5
    x1 = mesh.node_xyz
    x2 = O.motion(x1, mesh.node_disp, dt, 0.0, params["motion_omega"], params["motion_phase"])
    np.testing.assert_array_equal(xyz, x2)
    g1, g2 = O.measure(om, x1), O.measure(om, x2)
    assert np.abs(g2["cell_volume"] / g1["cell_volume"] - 1).max() > 1e-8   # the volumes change
    rc, flux, _, _ = O.face_flux(om, params, prim, g1, x1, x2, dt=dt)
    assert rc == 0
    rc, ref, _, _ = O.update(om, params, prim, g1["cell_volume"], g2["cell_volume"], flux, dt, src)
    assert rc == 0
    e = cell_errors(got, ref)
    print("fused surrogate, moving shell, precision %d: per-cell errors %s" % (precision, e))
    assert cell_err(got, ref) < (1e-12 if precision == 0 else 1e-6), e


def floor_state(C):
    """A cube whose left cells hold rho = 1e-8 (below the 1e-7 density floor after the update) and
    whose right cells hold p = 1e-10 (below the 1e-8 pressure floor)."""
    mesh, prim, params = C.cube_brio_wu(n=10)
    cc = mesh.node_xyz[mesh.cell_nodes].mean(axis=1)
    prim = prim.copy()
    lo, hi = cc[:, 0] < 0.3, cc[:, 0] > 0.6
    prim[lo, 0] = 1e-8
    prim[lo, 4] = 1e-9
    prim[lo, 5:8] = 0.0
    prim[hi, 4] = 1e-10
    return mesh, prim, params


@pytest.mark.parametrize("precision", [0, 1])
def test_floors_bit_exact(T, C, O, precision):
    """The density floor (rho' < 1e-7 -> 0.1, p = rho'/2, V and B kept) and the pressure floor
    (p' < 1e-8 -> 1e-6) fire in the same cells as the oracle's, with the same counts; FP64 bit for
    bit over 5 steps.  The mixed solver's fluxes differ in the last bits, yet the floors (exact
    constants) give the same states in the floored cells."""
    mesh, prim, params = floor_state(C)
    om = O.OracleMesh(mesh)
    g = O.measure(om, mesh.node_xyz)
    rc, flux, _, _ = O.face_flux(om, params, prim, g)
    rc, ref1, nr, npf = O.update(om, params, prim, g["cell_volume"], g["cell_volume"], flux, 1e-5)
    assert rc == 0 and nr > 0 and npf > 0
    with T.Setka(mesh, dict(params, precision=precision)) as s:
        s.set_state(prim, 0)
        st = s.Go(1)
        got1 = s.get_state()
        s.Go(4)
        got5 = s.get_state()
    assert (st["n_rho_floor"], st["n_p_floor"]) == (nr, npf)
    rho_floored = ref1[:, 0] == 0.1
    p_floored = ref1[:, 4] == 0.000001
    assert rho_floored.sum() == nr and p_floored.sum() == npf
    np.testing.assert_array_equal(got1[rho_floored], ref1[rho_floored])
    np.testing.assert_array_equal(got1[p_floored, 4], ref1[p_floored, 4])
    if precision == 0:
        np.testing.assert_array_equal(got1, ref1)
        rc, ref5, _, _, _ = O.go(om, params, mesh.node_xyz, prim, 5, nthreads=8)
        assert rc == 0
        np.testing.assert_array_equal(got5, ref5)


def test_failed_go_leaves_the_context_untouched(T, C):
    """Go with the exact source before any neutrals is TITAN_E_ARG and changes nothing: the layers,
    the step count and the lagged dt stay as they were, so after set_neutrals the first step is
    the same, bit for bit, as on a context that had its neutrals from the start."""
    mesh, prim, params = C.shell(n_edge=4, n_layers=6, noise=0.01)
    nh = C.neutrals(mesh.n_cells)
    p = dict(params, source_mode=T.SRC_EXACT, cx_nodes=16)
    with T.Setka(mesh, p) as s:
        s.set_state(prim, 0)
        with pytest.raises(T.TitanError) as e:
            s.Go(1)
        assert e.value.code == T.TITAN_E_ARG
        s.set_neutrals(nh)
        st = s.Go(2)
        got = s.get_state()
    with T.Setka(mesh, p) as s:
        s.set_state(prim, 0)
        s.set_neutrals(nh)
        st_ref = s.Go(2)
        ref = s.get_state()
    assert st["steps_done"] == st_ref["steps_done"] == 2
    assert st["dt_last"] == st_ref["dt_last"]
    np.testing.assert_array_equal(got, ref)
