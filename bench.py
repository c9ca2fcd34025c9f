#!/usr/bin/env python3
"""Benchmark: FP64 cell-updates/s of the zephyr finite-volume step on one MI355X.

A step is one SmFluid::update (sm_fluid.cpp:81-102) over the whole mesh:
dt reduction + gradient/limiting + predictor + corrector (MUSCL) with HLLC.
Inputs are resident in HBM before the timed region; the timed region is
exactly --steps steps bracketed by device synchronisations.

Prints one JSON line with the metric, a roofline object for the dominant
kernel (algorithmic bytes per launch / HIP-event-measured launch time) and a
CPU baseline (the oracle, OpenMP, on a bounded sample of the same workload).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK_GBS = 8000.0

# Algorithmic bytes per cell-update of each stage (phi = unique faces / cells),
# the stage-structured SoA accounting of the workload definition.
def stage_bytes(phi):
    return {
        "dt": 40.0,
        "flux1": 32 + 8 + 32 * phi + 32,
        "grad": 32 + 16 + 32 * phi + 64,
        "pred": 32 + 64 + 4 + 16 + 8 + 40 * phi + 32,
        "corr": 32 + 32 + 64 + 4 + 16 + 8 + 48 * phi + 32 + 4,
    }


def kernel_bytes_per_cell(phi):
    s = stage_bytes(phi)
    return {
        "flux1": s["flux1"] + s["dt"],          # first-order update + fused next-step dt
        "grad_pred": s["grad"] + s["pred"],      # gradient + limiting + predictor (fused)
        "corr": s["corr"] + s["dt"],             # corrector + fused next-step dt
        "dt": s["dt"],
    }


WORKLOADS = {
    # name: (nx, ny, acc, limiter, description)
    "uniform1024_muscl": (1024, 1024, 2, "MC",
                          "1024x1024 uniform mesh, MUSCL-Hancock + MC + HLLC, FP64, ZOE bounds, "
                          "seeded Gaussian random density field (mean 1, std 0.1, corr. length 0.05), p = 1"),
    "quadrants200_first_order": (200, 200, 1, "MC",
                                 "200x200 mesh, first-order HLLC, CFL 0.5, seeded four-quadrant Riemann data"),
}


def initial_state(name, nx, ny, seed):
    from zephyr_amd import problems
    if name.startswith("uniform"):
        return problems.smooth_density_field(nx, ny, seed)
    i, j = np.divmod(np.arange(nx * ny), ny)
    cx = (i + 0.5) / nx
    cy = (j + 0.5) / ny
    return problems.four_quadrant(cx, cy, seed)


def cpu_baseline(name, nx, ny, acc, limiter, seed, budget_s):
    """Oracle (CPU restatement, OpenMP) on the same mesh and scheme for as many
    steps as fit in budget_s (at least 2)."""
    from oracle import pyoracle as O
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(16, os.cpu_count() or 1)
    m = O.Mesh.rectangle(nx, ny)
    s = O.SmFluid(m, acc=acc, flux=O.HLLC, limiter=limiter, cfl=0.5)
    s.set_state(*initial_state(name, nx, ny, seed))
    s.update(threads)   # untimed warm-up step
    steps, t0 = 0, time.perf_counter()
    while True:
        s.update(threads)
        steps += 1
        el = time.perf_counter() - t0
        if steps >= 2 and el >= budget_s:
            break
    return {"value": nx * ny * steps / el, "unit": "cell-updates/s", "cores": threads, "kind": "port",
            "sample": f"oracle (C++ -O3 -ffp-contract=off, OpenMP {threads} threads) on the same "
                      f"{nx}x{ny} workload, {steps} steps in {el:.2f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="uniform1024_muscl", choices=sorted(WORKLOADS))
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--cpu-budget", type=float, default=10.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run device-to-device copies of the state (known bytes) for HBM counter scaling")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only; --gpus {args.gpus} is not supported "
              f"(the sharded design is described in DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import zephyr_amd as Z

    nx, ny, acc, limiter, desc = WORKLOADS[args.workload]
    n = nx * ny
    ctx = Z.Context()
    ctx.mesh_rectangle(nx, ny)
    ctx.configure(acc=acc, flux="hllc", limiter=limiter, cfl=0.5, gamma=1.4)
    ctx.set_state(*initial_state(args.workload, nx, ny, args.seed))

    ctx.step(args.warmup)
    ctx.synchronize()

    if args.copy_calibration:
        # 5 device-to-device copies of n doubles (the init layer) into fresh
        # buffers: 8n bytes read and 8n bytes written per copy dispatch.
        import ctypes as C
        hip = C.CDLL("libamdhip64.so")
        bufs = [C.c_void_p() for _ in range(5)]
        for b in bufs:
            if hip.hipMalloc(C.byref(b), C.c_size_t(8 * n)) != 0:
                raise RuntimeError("hipMalloc failed")
        for _ in range(3):
            Z._check(Z.lib().zfv_state_download_device(ctx.h, *bufs))
        for b in bufs:
            hip.hipFree(b)

    ctx.profile_enable(True)
    ctx.synchronize()
    t0 = time.perf_counter()
    ctx.step(args.steps)
    ctx.synchronize()
    el = time.perf_counter() - t0
    prof = ctx.profile_read()
    ctx.profile_enable(False)

    value = n * args.steps / el
    phi = 2.0 + (nx + ny) / float(n)
    kb = kernel_bytes_per_cell(phi)
    per_kernel = {k: (ms / max(cnt, 1)) for k, (ms, cnt) in prof.items()}
    dom = max((k for k in prof if k != "dt"), key=lambda k: prof[k][0])
    dom_ms = per_kernel[dom]
    alg_bytes = kb[dom] * n
    achieved = alg_bytes / (dom_ms * 1e-3) / 1e9
    step_bytes = sum(kb[k] for k in (("flux1",) if acc == 1 else ("grad_pred", "corr"))) * n

    traffic = None
    tpath = os.path.join(HERE, "profiles", "traffic.json")
    if os.path.exists(tpath):
        t = json.load(open(tpath)).get(args.workload, {}).get(dom)
        if t:
            traffic = t

    out = {
        "metric": "FP64 cell-updates/s (HLLC+MUSCL)" if acc == 2 else "FP64 cell-updates/s (HLLC first order)",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": args.workload, "description": desc, "cells": n, "phi": phi,
                   "scheme": "MUSCL+HLLC" if acc == 2 else "first-order HLLC", "limiter": limiter,
                   "step_algorithmic_bytes": step_bytes,
                   "step_roofline_frac": step_bytes / (el / args.steps) / 1e9 / HBM_PEAK_GBS,
                   "kernel_ms": per_kernel},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "algorithmic_bytes_per_launch": alg_bytes},
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.workload, nx, ny, acc, limiter, args.seed, args.cpu_budget)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
