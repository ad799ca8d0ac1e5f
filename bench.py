"""Benchmark of the explicit ideal-MHD step on one MI355X.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload shell4m_moving] [--no-cpu]
                [--precision-steps 1000] [--train-states 2000000] [--exact-steps 3] [--variants ...]

Prints ONE JSON line.  `value` is cell-updates/s (N_c x K / wall time of exactly K timed steps,
inputs resident in HBM) of the plain FP64 step, the variant whose kernel the HBM roofline is
quoted on.  Beside it:
  variants    cell-updates/s of (i) plain FP64, FP32-mixed, (ii) FP64 + exact charge-exchange
              source (FP64 Maxwellian quadrature every step), (iii) FP32-mixed + surrogate source
              (MLP fused into the step kernel, exact fallback for out-of-box cells), and the
              ratios (iii)/(ii) and (iii)/(i); FP64 / FP32-mixed with second-order (TVD)
              face states (Setka_gran_snos.cpp:186-736); FP64 carrying the passive scalars Q and
              rho_He (which the reference's step always carries); FP64 with the reference's
              closed-form source (Calc_sourse_MF_Bera) and four hydrogen fluids;
  checkpoint  save and load of the state and nodes in the reference's cell-parameter file;
  roofline    the step kernel of (i): algorithmic bytes per launch / its mean duration, timed with
              HIP events around every launch on the context's stream;
  surrogate   training-set size, epochs, loss, and its source error against the exact kernel;
  precision   FP32-mixed vs FP64 after --precision-steps steps from the same seeded input: per-cell
              relative errors (largest, where, percentiles), field L2/Linf and the div B norms;
  accuracy_gate  (iii) FP32-mixed + surrogate vs (ii) FP64 + exact source after --gate-steps steps
              on --gate-workload (default: the 1M-cell moving shell), per cell;
  shell1m_sizes  plain FP64 on the 1M-cell fixed and moving shells with their own roofline;
  cpu_baseline  the oracle's FP64 step (OpenMP over faces and cells) on this host's cores, on a
              bounded sample of the same workload.
"""
import argparse
import importlib
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
# speed bins and range of the pickup-ion source tallies timed beside the Monte-Carlo call
PUI_BINS, PUI_RANGE = 50, 20.0
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_k_step_tile_fp64_%s.md")
VALU_FILE = os.path.join(ROOT, "profiles", "valu_k_step_tile_fp64_%s.md")
# MI355X VALU issue: 256 CUs x 4 SIMDs, one wave64 instruction per 4 cycles per SIMD at 2.4 GHz
VALU_PEAK_WAVE_INSTR_PER_S = 256 * 4 * 2.4e9 / 4


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def algorithmic_bytes(n_cells, n_faces, n_nodes, moving, with_neutrals=False):
    """Compulsory HBM traffic per step (design independent): state read + write 128 B/cell,
    cell volume (fixed mesh) or 8 node ids (moving) per cell, per face the two cell ids plus the
    unit normal and area (fixed) or 4 node ids (moving), per node old + new xyz (moving);
    neutral moments 40 B/cell for the charge-exchange source."""
    if moving:
        b = 160 * n_cells + 24 * n_faces + 48 * n_nodes
    else:
        b = 136 * n_cells + 40 * n_faces
    if with_neutrals:
        b += 40 * n_cells
    return float(b)


def run_variant(T, mesh, prim, params, steps, warmup, neutrals=None, mlp=None, keep_state=False, h_fluids=None,
                rays=None, scal=None):
    s = T.Setka(mesh, params)
    try:
        s.set_state(prim, 0)
        if scal is not None:  # the passive scalars Q, rho_He (params passive_scalars = 1)
            s.set_scalars(scal, 0)
        if rays is not None:  # fitted surfaces: rays and the smoothing set-up of TITAN.cpp:108-112
            s.set_rays(rays)
            s.Smooth_head_TS3()
            s.Smooth_head_HP3()
            s.Find_Yzel_Sosed_for_BS()
        if neutrals is not None:
            s.set_neutrals(neutrals)
        if h_fluids is not None:
            s.set_h_fluids(h_fluids)
        if mlp is not None:
            s.mlp_set(*mlp)
        if warmup:
            s.Go(warmup)
        s.set_timing(True)
        t0 = time.perf_counter()
        st = s.Go(steps)  # returns after the stream has drained
        t1 = time.perf_counter()
        out = dict(seconds=t1 - t0, steps=steps, stats=st, step_kernel_ms=s.kernel_ms(0),
                   source_kernel_ms=s.kernel_ms(3), motion_kernel_ms=s.kernel_ms(2), tiles=s.tile_info())
        if keep_state:
            out["state"] = s.get_state()
        return out
    finally:
        s.close()


def divb(T, mesh, params, prim, xyz):
    """div B per cell, (1/V) sum_f s_f PdivB_f with PdivB = 1/2 (B_L + B_R).n A (the face loop's
    ninth flux component), on the given state and node positions."""
    with T.Setka(mesh, dict(params, move_mesh=0)) as s:
        s.set_nodes(xyz, 0)
        s.set_state(prim, 0)
        flux, _, _ = s.Culc_Gran_Potok(0, 1e-5, want_branch=False)
        vol = s.geometry(0)["cell_volume"]
    cf = mesh.cell_faces
    sign = np.where(mesh.face_cells[cf, 0] == np.arange(mesh.n_cells)[:, None], 1.0, -1.0)
    return (sign * flux[cf, 8]).sum(axis=1) / vol, vol


GAMMA = 5.0 / 3.0
AU = 0.237455   # 1 AU in code units (Phys_param.h:117)


def per_cell_errors(got, ref, r_au):
    """Per-cell relative errors of `got` against `ref`, each cell against its own value: rho and p
    by |d|/|ref|, the total energy E likewise, momentum rho v and B by the norm of the vector
    difference over the cell's |rho v| or |B|.  For each: the largest value, the radius (AU) of the
    cell that has it, and the 99th / 99.9th percentiles and the median over the cells."""
    def cons(q):
        E = q[:, 4] / (GAMMA - 1) + 0.5 * q[:, 0] * (q[:, 1:4] ** 2).sum(1) + (q[:, 5:8] ** 2).sum(1) / (8 * np.pi)
        return q[:, 0:1] * q[:, 1:4], E
    mg, Eg = cons(got)
    mr, Er = cons(ref)
    tiny = 1e-300
    e = {"rho": np.abs(got[:, 0] - ref[:, 0]) / np.maximum(np.abs(ref[:, 0]), tiny),
         "p": np.abs(got[:, 4] - ref[:, 4]) / np.maximum(np.abs(ref[:, 4]), tiny),
         "E": np.abs(Eg - Er) / np.maximum(np.abs(Er), tiny),
         "rho_v": np.linalg.norm(mg - mr, axis=1) / np.maximum(np.linalg.norm(mr, axis=1), tiny),
         "B": np.linalg.norm(got[:, 5:8] - ref[:, 5:8], axis=1) / np.maximum(np.linalg.norm(ref[:, 5:8], axis=1), tiny)}
    out = {}
    for k, v in e.items():
        i = int(np.argmax(v))
        out[k] = {"max": float(v[i]), "r_au_at_max": round(float(r_au[i]), 2),
                  "p99": float(np.percentile(v, 99)), "p999": float(np.percentile(v, 99.9)),
                  "median": float(np.median(v))}
    return out


def cell_radius_au(mesh):
    return np.linalg.norm(mesh.node_xyz[mesh.cell_nodes].mean(axis=1), axis=1) / AU


def ulp_perturbed(prim, seed=7):
    """The state with every value moved by one unit in the last place, up or down (seeded)."""
    rng = np.random.default_rng(seed)
    return prim * (1.0 + rng.choice([-1.0, 1.0], prim.shape) * 2.0 ** -52)


def precision_report(T, mesh, prim, params, steps):
    """FP32-mixed vs FP64 from the same input; beside it, FP64 from the input moved by one ulp vs
    FP64: how far the dynamics carry a last-bit difference in the same number of steps, the floor
    below which no arithmetic other than the FP64 step's own can agree per cell."""
    res = {}
    for name, prec, p0 in (("fp64", T.PREC_FP64, prim), ("mixed", T.PREC_MIXED, prim),
                           ("fp64_ulp", T.PREC_FP64, ulp_perturbed(prim))):
        with T.Setka(mesh, dict(params, precision=prec)) as s:
            s.set_state(p0, 0)
            t0 = time.perf_counter()
            st = s.Go(steps)
            res[name] = dict(state=s.get_state(), xyz=s.get_nodes(s.current_layer()), stats=st,
                             seconds=time.perf_counter() - t0)
    a, b = res["mixed"]["state"], res["fp64"]["state"]
    groups = {"rho": [0], "V": [1, 2, 3], "p": [4], "B": [5, 6, 7]}
    out = {"steps": steps, "reference": "FP64 GPU step (bit-exact with the oracle at the parity sizes)",
           "norm": "per_cell: |d|/|ref| per cell (vectors: norm of the difference over the cell's own norm); "
                   "field: rel_l2 over the mesh and rel_linf = max|d| / max|ref| of the field"}
    out["per_cell"] = per_cell_errors(a, b, cell_radius_au(mesh))
    out["ulp_sensitivity_per_cell"] = per_cell_errors(res["fp64_ulp"]["state"], b, cell_radius_au(mesh))
    out["field"] = {}
    for g, cols in groups.items():
        d, r = a[:, cols] - b[:, cols], b[:, cols]
        out["field"][g] = {"rel_l2": float(np.linalg.norm(d) / np.linalg.norm(r)),
                           "rel_linf": float(np.abs(d).max() / np.abs(r).max())}
    bmag = np.linalg.norm(b[:, 5:8], axis=1)
    for name in ("fp64", "mixed"):
        dv, vol = divb(T, mesh, params, res[name]["state"], res[name]["xyz"])
        h = np.cbrt(vol)
        out["divB_" + name] = {"l2_h_over_B": float(np.sqrt(np.mean((dv * h / np.maximum(bmag, 1e-300)) ** 2))),
                               "max_abs": float(np.abs(dv).max())}
    out["rho_p_per_cell_within_1e-4"] = bool(max(out["per_cell"]["rho"]["max"], out["per_cell"]["p"]["max"]) <= 1e-4)
    out["mixed_fp64_face_solves_last_step"] = int(res["mixed"]["stats"]["n_fp64_faces"])
    out["t_sim"] = res["fp64"]["stats"]["t_sim"]
    return out


def gate_report(T, C, mlp, workload, steps):
    """The accuracy gate on the configuration it is about: (iii) FP32-mixed + surrogate source
    against (ii) FP64 + exact source (FP64 quadrature every step), `steps` steps from the same
    seeded input on `workload`, per cell.  Beside it, (i) FP64 without a source against (ii): how
    far the source itself moves the state, the scale the surrogate's error is to be read against."""
    mesh, prim, params = C.build(workload)
    nh = C.neutrals(mesh.n_cells)
    runs = {}
    for name, prec, mode, p0 in (("exact", T.PREC_FP64, T.SRC_EXACT, prim), ("surrogate", T.PREC_MIXED, T.SRC_MLP, prim),
                                 ("none", T.PREC_FP64, T.SRC_NONE, prim),
                                 ("none_ulp", T.PREC_FP64, T.SRC_NONE, ulp_perturbed(prim))):
        with T.Setka(mesh, dict(params, precision=prec, source_mode=mode, cx_nodes=100)) as s:
            s.set_state(p0, 0)
            s.set_neutrals(nh)
            s.mlp_set(*mlp)
            t0 = time.perf_counter()
            try:
                st = s.Go(steps)
            except T.TitanError as e:  # a run that stops with a status code is reported, not hidden
                log("gate %s stopped: %s" % (name, e))
                return {"workload": workload, "n_cells": mesh.n_cells, "steps": steps, "error": "%s: %s" % (name, e)}
            runs[name] = dict(state=s.get_state(), seconds=time.perf_counter() - t0, stats=st)
        log("gate %s: %d steps in %.1f s" % (name, steps, runs[name]["seconds"]))
    r_au = cell_radius_au(mesh)
    out = {"workload": workload, "n_cells": mesh.n_cells, "steps": steps,
           "compared": "(iii) FP32-mixed + surrogate vs (ii) FP64 + exact source, per cell",
           "per_cell": per_cell_errors(runs["surrogate"]["state"], runs["exact"]["state"], r_au),
           "source_effect_per_cell": per_cell_errors(runs["none"]["state"], runs["exact"]["state"], r_au),
           "ulp_sensitivity_per_cell": per_cell_errors(runs["none_ulp"]["state"], runs["none"]["state"], r_au),
           "fallback_cells_last_step": int(runs["surrogate"]["stats"]["n_fallback"]),
           "seconds": {k: round(v["seconds"], 2) for k, v in runs.items()}}
    pc = out["per_cell"]
    out["rho_p_per_cell_within_1e-4"] = bool(max(pc["rho"]["max"], pc["p"]["max"]) <= 1e-4)
    return out


def train_surrogate(T, C, sur, prim, nh, n_states, seed=11):
    """Training set: n_states plasma/neutral states drawn from the workload's cells (perturbed),
    labelled by the exact FP64 quadrature kernel; trained by the device's own forward/backward
    and Adam kernels."""
    t0 = time.perf_counter()
    in10 = sur.sample_states(prim, nh, n_states, seed=seed)
    x, y, _ = sur.make_dataset(in10, nodes=100)
    t_label = time.perf_counter() - t0
    norm = sur.Normaliser(x, y)
    epochs, lrs = (16, 10, 6), (2e-3, 5e-4, 1e-4)
    mesh_t, _, par_t = C.cube_brio_wu(n=4)
    with T.Setka(mesh_t, par_t) as s:
        t1 = time.perf_counter()
        hist = sur.train(s, x, y, norm, epochs=epochs, lrs=lrs, batch=4096)
        t_train = time.perf_counter() - t1
        P = s.mlp_get()
    in_norm, scale, box = norm.blocks()
    info = {"states": int(n_states), "label_s": round(t_label, 2), "train_s": round(t_train, 2),
            "epochs": int(sum(epochs)), "loss_first": float(hist[0]), "loss_last": float(hist[-1])}
    return (P, in_norm, scale, box), info


def checkpoint_timing(T, mesh, prim, params):
    """Save_cell_parameters / Download_cell_parameters of the whole state and node layer through a
    file in a temporary directory (device <-> host copies included)."""
    import tempfile
    with tempfile.TemporaryDirectory() as d, T.Setka(mesh, params) as s:
        path = os.path.join(d, "parameters.bin")
        s.set_state(prim, 0)
        t0 = time.perf_counter()
        s.Save_cell_parameters(path)
        t1 = time.perf_counter()
        nbytes = os.path.getsize(path)
        s.Download_cell_parameters(path)
        t2 = time.perf_counter()
    out = {"bytes": nbytes, "save_s": round(t1 - t0, 3), "load_s": round(t2 - t1, 3),
           "save_GBps": round(nbytes / (t1 - t0) / 1e9, 3), "load_GBps": round(nbytes / (t2 - t1) / 1e9, 3),
           "keys": 8, "nodes": True}
    log("checkpoint: %s" % out)
    return out


def traffic_from_profile(workload):
    """HBM bytes per launch of the FP64 step kernel measured by rocprofv3 FETCH_SIZE / WRITE_SIZE
    counter runs (calibrated on a copy kernel), as recorded in the profile summary, or None."""
    path = TRAFFIC_FILE % workload
    if not os.path.exists(path):
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", open(path).read())
    return float(m.group(1)) if m else None


def valu_from_profile(workload):
    """VALU wave-instructions per launch of the FP64 step kernel (rocprofv3 SQ_INSTS_VALU), as
    recorded in the profile summary, or None."""
    path = VALU_FILE % workload
    if not os.path.exists(path):
        return None
    m = re.search(r"valu_wave_instructions_per_launch:\s*([0-9.eE+]+)", open(path).read())
    return float(m.group(1)) if m else None


def copy_calibration():
    """A plain streaming kernel of known size (y = -x: 256 MiB read + 256 MiB written), for the
    counter runs that scale FETCH_SIZE / WRITE_SIZE.  Runs in a process of its own: torch ships its
    own HIP runtime, and a process holding both it and the one libtitan_gpu.so links is not one the
    counter collection handles."""
    import torch
    x = torch.ones(32 * 1024 * 1024, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        torch.neg(x, out=y)
    torch.cuda.synchronize()
    log("copy calibration: 3 x (%d bytes read + written)" % (x.numel() * 8))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="shell4m_moving")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=16)
    ap.add_argument("--exact-steps", type=int, default=3)
    ap.add_argument("--train-states", type=int, default=2_000_000)
    ap.add_argument("--precision-steps", type=int, default=1000)
    ap.add_argument("--gate-steps", type=int, default=1000,
                    help="steps of the (iii) mixed + surrogate vs (ii) FP64 + exact source comparison (0: skip)")
    ap.add_argument("--gate-workload", default="shell1m_moving")
    ap.add_argument("--variants",
                    default="fp64,mixed,fp64_exact,mixed_surrogate,fp64_tvd,mixed_tvd,fp64_scalars,fp64_bera,fp64_fit,mk,"
                            "checkpoint,shell1m,shell1m_moving")
    ap.add_argument("--mk-per-face", type=int, default=100,
                    help="Monte-Carlo neutrals: particles per outer-boundary face (N_per_gran)")
    ap.add_argument("--mk-cpu-per-face", type=int, default=100,
                    help="Monte-Carlo CPU baseline: particles per outer face (100: the same call as the GPU's)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run the streaming copy the HBM counter runs are scaled by, then exit")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py: this build runs on one GPU; runs on %d GPUs are not part of it (use --gpus 1)" % a.gpus,
              file=sys.stderr)
        sys.exit(2)
    if a.copy_calibration:
        copy_calibration()
        return
    T = importlib.import_module("titan_c-_amd")
    C = importlib.import_module("titan_c-_amd.configs")
    sur = importlib.import_module("titan_c-_amd.surrogate")
    T.lib()
    variants = a.variants.split(",")
    t0 = time.perf_counter()
    mesh, prim, params = C.build(a.workload)
    nh = C.neutrals(mesh.n_cells)
    moving = bool(params["move_mesh"])
    nc, nf, nn = mesh.n_cells, mesh.n_faces, mesh.n_nodes
    log("workload %s: %d cells, %d faces, %d nodes (built in %.1f s)" % (a.workload, nc, nf, nn,
                                                                         time.perf_counter() - t0))

    res = {}
    if "fp64" in variants:
        res["fp64"] = run_variant(T, mesh, prim, dict(params, precision=T.PREC_FP64), a.steps, a.warmup)
        log("fp64: %.3f ms/step, step kernel %.3f ms; tiles %s" % (
            res["fp64"]["seconds"] * 1e3 / a.steps, res["fp64"]["step_kernel_ms"], res["fp64"]["tiles"]))
    if "mixed" in variants:
        res["mixed"] = run_variant(T, mesh, prim, dict(params, precision=T.PREC_MIXED), a.steps, a.warmup)
        log("mixed: %.3f ms/step" % (res["mixed"]["seconds"] * 1e3 / a.steps))
    if "fp64_exact" in variants:
        p = dict(params, precision=T.PREC_FP64, source_mode=T.SRC_EXACT, cx_nodes=100)
        res["fp64_exact"] = run_variant(T, mesh, prim, p, a.exact_steps, 1, neutrals=nh)
        log("fp64 + exact source: %.1f ms/step (source kernel %.1f ms)" % (
            res["fp64_exact"]["seconds"] * 1e3 / a.exact_steps, res["fp64_exact"]["source_kernel_ms"]))
    sur_info = None
    if "mixed_surrogate" in variants:
        mlp, sur_info = train_surrogate(T, C, sur, prim, nh, a.train_states)
        log("surrogate trained: %s" % sur_info)
        with T.Setka(mesh, dict(params, cx_nodes=100)) as s:
            s.set_state(prim, 0)
            s.set_neutrals(nh)
            s.mlp_set(*mlp)
            ex = s.sources(0, T.SRC_EXACT)
            ml = s.sources(0, T.SRC_MLP)
        cols = ("rate", "Sm_x", "Sm_y", "Sm_z", "S_E")
        err = np.abs(ml - ex)
        sur_info["source_err_rel_max"] = {k: float(err[:, i].max() / np.abs(ex[:, i]).max())
                                          for i, k in enumerate(cols)}
        sur_info["source_err_rel_median"] = {
            k: float(np.median(err[:, i] / np.maximum(np.abs(ex[:, i]), 1e-300))) for i, k in enumerate(cols)}
        p = dict(params, precision=T.PREC_MIXED, source_mode=T.SRC_MLP, cx_nodes=100)
        res["mixed_surrogate"] = run_variant(T, mesh, prim, p, a.steps, a.warmup, neutrals=nh, mlp=mlp)
        sur_info["fallback_cells_last_step"] = int(res["mixed_surrogate"]["stats"]["n_fallback"])
        log("mixed + surrogate: %.3f ms/step" % (res["mixed_surrogate"]["seconds"] * 1e3 / a.steps))

    for name, prec in (("fp64_tvd", T.PREC_FP64), ("mixed_tvd", T.PREC_MIXED)):
        if name in variants:  # second-order (TVD) face states, first-order otherwise identical
            res[name] = run_variant(T, mesh, prim, dict(params, precision=prec, tvd=1), a.steps, a.warmup)
            log("%s: %.3f ms/step, step kernel %.3f ms, %d second-order faces" % (
                name, res[name]["seconds"] * 1e3 / a.steps, res[name]["step_kernel_ms"],
                res[name]["tiles"]["tvd_faces"]))

    if "fp64_scalars" in variants:  # the passive scalars Q and rho_He the reference always carries
        res["fp64_scalars"] = run_variant(T, mesh, prim, dict(params, precision=T.PREC_FP64, passive_scalars=1),
                                          a.steps, a.warmup, scal=C.passive_scalars(mesh, prim))
        log("fp64 + passive scalars: %.3f ms/step, step kernel %.3f ms" % (
            res["fp64_scalars"]["seconds"] * 1e3 / a.steps, res["fp64_scalars"]["step_kernel_ms"]))

    if "fp64_bera" in variants:  # the reference's closed-form charge-exchange source, four H fluids
        hf = C.h_fluids(nc)
        p = dict(params, precision=T.PREC_FP64, source_mode=T.SRC_BERA)
        res["fp64_bera"] = run_variant(T, mesh, prim, p, a.steps, a.warmup, h_fluids=hf)
        log("fp64 + closed-form source: %.3f ms/step (source kernel %.3f ms)" % (
            res["fp64_bera"]["seconds"] * 1e3 / a.steps, res["fp64_bera"]["source_kernel_ms"]))
        del hf
    fitted = C.build("shell4m_fitted") if ("fp64_fit" in variants or "mk" in variants) else None
    mk_info = None
    if "mk" in variants:  # Monte-Carlo neutral densities on the fitted shell's state
        fmesh, fprim, fparams, _ = fitted
        with T.Setka(fmesh, fparams) as s:
            s.set_state(fprim, 0)
            s.MK_go(1, seed=99)  # first launch: code load
            t0 = time.perf_counter()
            _, st = s.MK_go(a.mk_per_face, seed=1)
            sec = time.perf_counter() - t0
            # the same call with the pickup-ion source tallies (MK_source_S) and their finishing pass
            s.MK_source_S(PUI_BINS, PUI_RANGE)
            t0 = time.perf_counter()
            s.MK_go(a.mk_per_face, seed=1)
            psec = time.perf_counter() - t0
            s.MK_source_S(0, 0.0)
        mk_info = {"workload": "shell4m_fitted", "n_cells": fmesh.n_cells, "n_per_face": a.mk_per_face,
                   "stats": st, "seconds": sec, "particles_per_s": st["particles"] / sec,
                   "segments_per_s": st["segments"] / sec,
                   "with_pickup_sources": {"n_w": PUI_BINS, "w_r": PUI_RANGE, "seconds": psec}}
        log("MC neutrals with the pickup-ion tallies (%d bins): %.3f s" % (PUI_BINS, psec))
        log("MC neutrals: %d particles, %d segments, %d exchanges in %.3f s (%.3g segments/s)" % (
            st["particles"], st["segments"], st["exchanges"], sec, st["segments"] / sec))
        if not a.no_cpu:
            sys.path.insert(0, os.path.join(ROOT, "oracle"))
            import oracle as O
            om = O.OracleMesh(fmesh)
            geom = O.measure(om, fmesh.node_xyz)
            t0 = time.perf_counter()
            ncpu = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
            _, ost = O.mk_go(om, fmesh.node_xyz, fprim, fparams["v_inf"], fparams.get("par_a_2", 0.102046),
                             fparams.get("par_Kn", 40.0906), a.mk_cpu_per_face, seed=1, geom=geom, nthreads=ncpu)
            csec = time.perf_counter() - t0
            del om, geom
            mk_info["cpu_baseline"] = {
                "value": ost["segments"] / csec, "unit": "segments/s", "cores": ncpu, "kind": "port",
                "sample": "%d particles (%d per outer face) of the same shell4m_fitted flight, C restatement of "
                          "MK_fly_immit, OpenMP over particles" % (ost["particles"], a.mk_cpu_per_face)}
            log("MC cpu baseline: %.3g segments/s" % (ost["segments"] / csec))
    if "fp64_fit" in variants:  # the same cell count with the TS/HP/BS fitted and the rays moved
        fmesh, fprim, fparams, frays = fitted
        res["fp64_fit"] = run_variant(T, fmesh, fprim, dict(fparams, precision=T.PREC_FP64), a.steps, a.warmup,
                                      rays=frays)
        log("fp64 + fitted surfaces (%d cells): %.3f ms/step, step kernel %.3f ms, surface+ray kernels %.3f ms" % (
            fmesh.n_cells, res["fp64_fit"]["seconds"] * 1e3 / a.steps, res["fp64_fit"]["step_kernel_ms"],
            res["fp64_fit"]["motion_kernel_ms"]))
        fit_info = {"workload": "shell4m_fitted", "n_cells": fmesh.n_cells,
                    "ms_per_step": res["fp64_fit"]["seconds"] * 1e3 / a.steps,
                    "step_kernel_ms": res["fp64_fit"]["step_kernel_ms"],
                    "surface_and_ray_kernels_ms": res["fp64_fit"]["motion_kernel_ms"],
                    "tiles": res["fp64_fit"]["tiles"]}
        del fmesh, fprim
    else:
        fit_info = None
    del fitted
    ckpt = checkpoint_timing(T, mesh, prim, params) if "checkpoint" in variants else None
    sizes = {}
    for wl in ("shell1m", "shell1m_moving"):  # the 1M-cell fixed and moving shells, plain FP64
        if wl in variants:
            m1, p1, par1 = C.build(wl)
            r = run_variant(T, m1, p1, dict(par1, precision=T.PREC_FP64), a.steps, a.warmup)
            mv = bool(par1["move_mesh"])
            bts = algorithmic_bytes(m1.n_cells, m1.n_faces, m1.n_nodes, mv)
            sizes[wl] = {"n_cells": m1.n_cells, "n_faces": m1.n_faces, "moving": mv,
                         "cell_updates_per_s": m1.n_cells * r["steps"] / r["seconds"],
                         "ms_per_step": r["seconds"] * 1e3 / r["steps"], "step_kernel_ms": r["step_kernel_ms"],
                         "algorithmic_bytes": bts,
                         "hbm_frac": round(bts / (r["step_kernel_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
            log("%s: %s" % (wl, sizes[wl]))
            del m1, p1
    gate = None
    if a.gate_steps > 0 and "mixed_surrogate" in variants and sur_info is not None:
        gate = gate_report(T, C, mlp, a.gate_workload, a.gate_steps)
        if "per_cell" in gate:
            log("gate: per-cell rho %s p %s" % (gate["per_cell"]["rho"], gate["per_cell"]["p"]))

    def cups(name):
        r = res.get(name)
        return None if r is None else nc * r["steps"] / r["seconds"]

    vals = {k: cups(k) for k in ("fp64", "mixed", "fp64_exact", "mixed_surrogate", "fp64_tvd", "mixed_tvd",
                                 "fp64_scalars", "fp64_bera", "fp64_fit")}
    ratios = {}
    if vals["mixed_surrogate"] and vals["fp64_exact"]:
        ratios["surrogate_vs_exact_source"] = vals["mixed_surrogate"] / vals["fp64_exact"]
    if vals["mixed_surrogate"] and vals["fp64"]:
        ratios["surrogate_vs_plain_fp64"] = vals["mixed_surrogate"] / vals["fp64"]
    if vals["mixed"] and vals["fp64"]:
        ratios["mixed_vs_plain_fp64"] = vals["mixed"] / vals["fp64"]
    if vals["fp64_tvd"] and vals["fp64"]:
        ratios["tvd_vs_first_order_fp64"] = vals["fp64_tvd"] / vals["fp64"]
    if vals["fp64_scalars"] and vals["fp64"]:
        ratios["passive_scalars_vs_plain_fp64"] = vals["fp64_scalars"] / vals["fp64"]

    roof = None
    head = "fp64" if "fp64" in res else None
    if head:
        bytes_step = algorithmic_bytes(nc, nf, nn, moving)
        k_ms = res[head]["step_kernel_ms"]
        ach = bytes_step / (k_ms * 1e-3) / 1e9
        roof = {"bound": "hbm", "kernel": "k_step_tile<double,false,false,false> (one launch = one step)",
                "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                "traffic": traffic_from_profile(a.workload), "traffic_unit": "bytes/launch",
                "algorithmic_bytes": bytes_step, "kernel_ms": round(k_ms, 4)}
        nv = valu_from_profile(a.workload)
        if nv:  # the roofline that binds: VALU issue (SQ_INSTS_VALU of the same kernel and workload)
            va = nv / (k_ms * 1e-3)
            roof["valu"] = {"achieved": round(va / 1e9, 1), "peak": round(VALU_PEAK_WAVE_INSTR_PER_S / 1e9, 1),
                            "unit": "G wave-instructions/s", "frac": round(va / VALU_PEAK_WAVE_INSTR_PER_S, 4),
                            "instructions_per_launch": nv}
            roof["binding"] = "valu"

    prec = None
    if a.precision_steps > 0:
        try:
            prec = precision_report(T, mesh, prim, params, a.precision_steps)
            log("precision after %d steps: per-cell rho %s p %s" % (a.precision_steps, prec["per_cell"]["rho"],
                                                                     prec["per_cell"]["p"]))
        except T.TitanError as e:  # a step that stops with a status code is reported, not hidden
            prec = {"steps": a.precision_steps, "error": str(e)}
            log("precision run stopped: %s" % e)

    cpu = None
    if not a.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        ncpu = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        om = O.OracleMesh(mesh)
        t0 = time.perf_counter()
        rc, _, _, _, _ = O.go(om, dict(params, precision=0), mesh.node_xyz, prim, a.cpu_steps, nthreads=ncpu,
                              disp=mesh.node_disp)
        t1 = time.perf_counter()
        cpu = {"value": nc * a.cpu_steps / (t1 - t0), "unit": "cell-updates/s", "cores": ncpu, "kind": "port",
               "sample": "%d FP64 steps of the same %s workload, first-order HLLC/HLLD, moving mesh re-measured "
                         "every step (C restatement of Setka::Go, OpenMP face and cell loops)" % (
                             a.cpu_steps, a.workload)}
        log("cpu baseline: %.3g cell-updates/s on %d threads (rc %d)" % (cpu["value"], ncpu, rc))

    r0 = res.get(head) if head else None
    out = {
        "metric": "cell-updates/s @4M cells: FP64 plain vs FP32-mixed+surrogate; % HBM peak",
        "value": vals.get("fp64"), "unit": "cell-updates/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": (r0["seconds"] * 1e3 / r0["steps"]) if r0 else None, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": a.workload, "n_cells": nc, "n_faces": nf, "n_nodes": nn, "moving": moving,
                   "variant": "plain FP64 step"},
        "variants": vals,
        "ratios": ratios,
        "roofline": roof,
        "surrogate": sur_info,
        "precision": prec,
        "accuracy_gate": gate,
        "shell1m_sizes": sizes,
        "checkpoint": ckpt,
        "fitted_surfaces": fit_info,
        "mc_neutrals": mk_info,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
