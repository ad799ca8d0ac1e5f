"""Benchmark of the explicit ideal-MHD step on one MI355X.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload shell4m_moving] [--no-cpu]

Prints ONE JSON line.  `value` is cell-updates/s (N_c x steps / wall time of exactly K timed
steps, inputs resident in HBM) of the headline variant; the other variants, the roofline of the
step kernel (HIP-event timing on the context's stream) and the CPU baseline (the oracle's
OpenMP FP64 step on this host's cores, bounded sample) are reported beside it.
"""
import argparse
import importlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


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


def run_variant(T, mesh, prim, params, steps, warmup, neutrals=None, mlp=None):
    s = T.Setka(mesh, params)
    try:
        s.set_state(prim, 0)
        if neutrals is not None:
            s.set_neutrals(neutrals)
        if mlp is not None:
            s.mlp_set(*mlp)
        if warmup:
            s.Go(warmup)
        s.set_timing(True)
        t0 = time.perf_counter()
        st = s.Go(steps)
        t1 = time.perf_counter()
        kms = {k: s.kernel_ms(i) for i, k in enumerate(("step", "geometry", "motion", "source"))}
        return dict(seconds=t1 - t0, stats=st, kernel_ms=kms, state=s.get_state(), tiles=s.tile_info())
    finally:
        s.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="shell4m_moving")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=2)
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py: this build targets one GPU; multi-GPU runs are not part of it (use --gpus 1)",
              file=sys.stderr)
        sys.exit(2)
    T = importlib.import_module("titan_c-_amd")
    C = importlib.import_module("titan_c-_amd.configs")
    T.lib()
    t0 = time.perf_counter()
    mesh, prim, params = C.build(a.workload)
    t_build = time.perf_counter() - t0
    moving = bool(params["move_mesh"])
    nc, nf, nn = mesh.n_cells, mesh.n_faces, mesh.n_nodes
    print("[bench] workload %s: %d cells, %d faces, %d nodes (built in %.1f s)" % (a.workload, nc, nf, nn, t_build),
          file=sys.stderr, flush=True)

    res = {}
    p64 = dict(params, precision=T.PREC_FP64)
    res["fp64"] = run_variant(T, mesh, prim, p64, a.steps, a.warmup)
    print("[bench] fp64 done: %.3f s for %d steps; tiles %s" % (res["fp64"]["seconds"], a.steps, res["fp64"]["tiles"]),
          file=sys.stderr, flush=True)
    pmx = dict(params, precision=T.PREC_MIXED)
    res["mixed"] = run_variant(T, mesh, prim, pmx, a.steps, a.warmup)
    print("[bench] mixed done: %.3f s" % res["mixed"]["seconds"], file=sys.stderr, flush=True)

    def cups(r):
        return nc * a.steps / r["seconds"]

    step_ms = res["fp64"]["seconds"] * 1e3 / a.steps
    bytes_step = algorithmic_bytes(nc, nf, nn, moving)
    # dominant kernel family: the face flux + cell update kernels of a step
    k_ms = res["fp64"]["kernel_ms"]["step"]
    ach = bytes_step / (step_ms * 1e-3) / 1e9
    roof = {"bound": "hbm", "kernel": "full step (all kernels of one step)", "achieved": round(ach, 1),
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None,
            "bytes_per_step": bytes_step, "step_kernels_ms": k_ms}

    cpu = None
    if not a.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        ncpu = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        om = O.OracleMesh(mesh)
        t0 = time.perf_counter()
        rc, _, _, _, _ = O.go(om, p64, mesh.node_xyz, prim, a.cpu_steps, nthreads=ncpu,
                              disp=mesh.node_disp)
        t1 = time.perf_counter()
        cpu = {"value": nc * a.cpu_steps / (t1 - t0), "unit": "cell-updates/s", "cores": ncpu, "kind": "port",
               "sample": "%d FP64 steps of the same %s workload (oracle, OpenMP face+cell loops, geometry "
                         "re-measured each step)" % (a.cpu_steps, a.workload)}

    out = {
        "metric": "cell-updates/s @4M cells: FP64 plain vs FP32-mixed+surrogate; % HBM peak",
        "value": cups(res["fp64"]), "unit": "cell-updates/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": step_ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": a.workload, "n_cells": nc, "n_faces": nf, "n_nodes": nn, "moving": moving,
                   "variant": "plain FP64 step"},
        "variants": {"fp64": cups(res["fp64"]), "mixed_fp32_riemann": cups(res["mixed"])},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
