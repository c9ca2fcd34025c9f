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
RESERVE_PER_BASE_CELL = 4   # AMR workloads: reserved cells per base cell

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
        "tile": s["grad"] + s["pred"] + s["corr"] + s["dt"],   # the whole MUSCL step in one tiled kernel
        "dt": s["dt"],
    }


# Bytes per cell the tiled MUSCL step itself must move (its fused minimum):
# the state in (rho, u, v, p, e) and out, the isolation counter in and out,
# volume and incircle; plus the halo of its tile shape (rows x 62 cells, a
# two-cell ring of rho, u, v, p and a one-cell ring of e and the counter),
# which neighbouring tiles read again.
TILE_ROWS = {512: 6, 768: 10, 1024: 14}


def tile_fused_bytes_per_cell(threads):
    rows = TILE_ROWS[threads]
    out = rows * 62.0
    ring2 = (rows + 4) * 66.0 / out - 1.0
    ring1 = (rows + 2) * 64.0 / out - 1.0
    return 40 + 4 + 8 + 8 + 40 + 4 + 32 * ring2 + 12 * ring1


WORKLOADS = {
    "amr256_muscl": dict(
        kind="amr", nx=256, ny=256, acc=2, limiter="minmod", max_level=3, threshold=0.1, every=4,
        desc="base 256x256 AMR-2D mesh, up to 3 refinement levels, MUSCL-Hancock + minmod + HLLC, CFL 0.5, FP64, "
             "ZOE bounds; seeded circle (radius 0.2, rho = p = 1 inside, 0.1 outside, 1% sinusoidal density "
             "perturbation); flags |grad rho| h / rho > 0.1 (hysteresis 0.4) and slope-distributor "
             "refine/coarsen every 4 steps, inside the timed region"),
    "uniform1024_muscl": dict(
        kind="uniform", nx=1024, ny=1024, acc=2, limiter="MC",
        desc="1024x1024 uniform mesh, MUSCL-Hancock + MC + HLLC, FP64, ZOE bounds, seeded Gaussian random "
             "density field (mean 1, std 0.1, corr. length 0.05), p = 1"),
    "quadrants200_first_order": dict(
        kind="uniform", nx=200, ny=200, acc=1, limiter="MC",
        desc="200x200 mesh, first-order HLLC, CFL 0.5, seeded four-quadrant Riemann data"),
    "adjoint1024_muscl": dict(
        kind="adjoint", nx=1024, ny=1024, acc=2, limiter="MC", every=20, window=(0.4, 0.6, 0.4, 0.6),
        steps=200, fd_dirs=64, fd_eps=(1e-6, 1e-7),
        desc="1024x1024 uniform mesh, MUSCL-Hancock + MC + HLLC, FP64, ZOE bounds; initial rho = 1 + 0.1 g0, "
             "p = 1 + 0.1 g1, u = 0.1 g2, v = 0.1 g3 (g_k: seeded unit Gaussian random fields, corr. length "
             "0.05); J = integral of p^2 "
             "over [0.4,0.6]^2 at the final step; gradient dJ/drho0 by the discrete adjoint (checkpoint every "
             "20 steps, recompute in between), checked against central differences (eps 1e-6 and 1e-7) on 64 seeded "
             "directions"),
    "mm512_bubble_first_order": dict(
        kind="mm", nx=512, ny=512, steps=300, cfl=0.5,
        desc="two-material shock-bubble on a 512x512 mesh (the C2 base mesh, without AMR): stiffened-gas water "
             "(gamma 4.4, P0 6e8) with an air bubble (gamma 1.4) of radius 0.15 at the centre, a planar shock of "
             "seeded Mach number 1.5-3 at x = 0.1 moving into the water; MmFluid first order with its defaults: "
             "material-aware HLLC, composite Riemann fluxes at faces next to mixed cells with the PLIC face "
             "fraction (interface reconstruction every step), PT-equilibrium closure; CFL 0.5, FP64, ZOE bounds"),
    "mm512_bubble_amr2": dict(
        kind="mm", nx=512, ny=512, steps=300, cfl=0.5, max_level=2, every=4,
        desc="the two-material shock-bubble of mm512_bubble_first_order on the C2 configuration: 512x512 base, up "
             "to 2 refinement levels by MmFluid::set_flags (mixed cells and cells next to another material) and "
             "the MmFluid split / merge distributor every 4 steps inside the timed region; initial adaptation in "
             "max_level + 1 rounds with the initial data re-sampled on each mesh"),
    "mm512_bubble_second_order": dict(
        kind="mm", nx=512, ny=512, steps=300, cfl=0.5, acc=2,
        desc="the two-material shock-bubble of mm512_bubble_first_order with MmFluid's second order "
             "(set_accuracy(2)): limited LSM gradients of every PState component (MC, zero in and next to mixed "
             "cells), the half-step predictor on pure cells, and the corrector on face states extrapolated from "
             "the half step and re-closed by the PT closure (interpolation_update); composite fluxes wherever an "
             "extrapolated face state is mixed"),
    "ensemble256_muscl": dict(
        kind="ensemble", problems=256, smin=64, smax=512, acc=2, limiter="MC", steps=200,
        desc="256 independent problems of s x s cells, s ~ U{64..512} (seeded), per-problem seeded "
             "four-quadrant Riemann data (seed + problem index), MUSCL-Hancock + MC + HLLC, FP64, CFL 0.5, "
             "per-problem dt; one context, one launch per stage for all problems, captured in HIP graphs, "
             "launch configuration autotuned; compared with the plain per-problem path looped over 256 contexts"),
    "cube128_muscl": dict(
        kind="cube", nx=128, ny=128, nz=128, acc=2, limiter="MC", steps=100,
        desc="128x128x128 Cuboid mesh (AMR-3D cells, one level), MUSCL-Hancock + MC + HLLC, CFL 0.5, FP64, ZOE "
             "bounds; seeded spherical blast (radius 0.25: p = 10, rho = 1.5 inside; p = rho = 1 outside) with "
             "a 1% seeded sinusoidal density perturbation and a smooth velocity field of amplitude 0.05"),
}


# The sources a step kernel is compiled from: its own file and the headers
# that file includes.
STEP_HEADERS = ("zfv_internal.hpp", "zfv_physics.hpp", "zfv_mm.hpp", "zfv_step_common.hpp", "zfv_step_faces.hpp")
KERNEL_FILES = {"tile": "zfv_tile.hip", "grad_pred": "zfv_kernels.hip", "corr": "zfv_kernels.hip",
                "flux1": "zfv_kernels.hip", "k3_grad_pred": "zfv_3d.hip", "k3_corr": "zfv_3d.hip",
                "k3_flux1": "zfv_3d.hip"}
KERNEL_HEADERS = {"k3_grad_pred": ("zfv_3d.hpp",), "k3_corr": ("zfv_3d.hpp",), "k3_flux1": ("zfv_3d.hpp",)}


def kernel_sources_sha(kernel):
    """Short SHA-256 over the HIP sources a step kernel is compiled from
    (which the HBM counter summaries record, so a summary is used only for
    the kernel code it measured); None for a kernel outside the table."""
    import hashlib
    if kernel not in KERNEL_FILES:
        return None
    d = os.path.join(HERE, "zephyr_amd", "csrc")
    h = hashlib.sha256()
    for f in (KERNEL_FILES[kernel],) + STEP_HEADERS + KERNEL_HEADERS.get(kernel, ()):
        h.update(f.encode())
        h.update(open(os.path.join(d, f), "rb").read())
    return h.hexdigest()[:16]


def ensemble_sizes(w, seed):
    """Side lengths s_i ~ U{smin..smax} from SplitMix64 (draw i -> problem i)."""
    from zephyr_amd import problems
    g = problems.SplitMix64(seed)
    span = w["smax"] - w["smin"] + 1
    return [w["smin"] + int(span * g.uniform()) for _ in range(w["problems"])]


def ensemble_state(s, seed, p):
    cx, cy = rect_centres(s, s)
    from zephyr_amd import problems
    return problems.four_quadrant(cx, cy, seed + p)


def initial_state(name, cx, cy, seed):
    from zephyr_amd import problems
    w = WORKLOADS[name]
    if w["kind"] == "amr":
        return problems.circle_blast(cx, cy, seed)
    if name.startswith("uniform"):
        return problems.smooth_density_field(w["nx"], w["ny"], seed)
    return problems.four_quadrant(cx, cy, seed)


def rect_centres(nx, ny):
    i, j = np.divmod(np.arange(nx * ny), ny)
    return (i + 0.5) / nx, (j + 0.5) / ny


def adapt_to_initial_data(solver_like, name, seed, centres, set_state, set_flags, refine, max_level, thr):
    """Initial adaptation: max_level + 1 rounds of (initial data at the cell
    centres, flags, refine with the const distributor), then the initial data
    on the final mesh."""
    for _ in range(max_level + 1):
        cx, cy = centres()
        set_state(*initial_state(name, cx, cy, seed))
        set_flags(1, thr)
        refine(max_level, False)
    cx, cy = centres()
    set_state(*initial_state(name, cx, cy, seed))


def host_info():
    """Host cores and CPU model of the machine the CPU baselines run on."""
    model = None
    try:
        import subprocess
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
                break
    except Exception:
        pass
    if model is None and os.path.exists("/proc/cpuinfo"):
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count()
    return {"nproc": os.cpu_count(), "affinity_cpus": affinity, "lscpu_model": model,
            "OMP_NUM_THREADS": os.environ.get("OMP_NUM_THREADS"), "OMP_PROC_BIND": os.environ.get("OMP_PROC_BIND")}


def cpu_threads():
    """All the host cores this process may use: OMP_NUM_THREADS when set (the
    CPU share of the machine), else the CPUs of the process's affinity mask."""
    env = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    if env > 0:
        return env
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_run(name, seed, budget_s, threads):
    """Oracle (CPU restatement, OpenMP) on the same workload for as many steps
    as fit in budget_s (at least 2); returns (cell-updates/s, steps, seconds)."""
    from oracle import pyoracle as O
    w = WORKLOADS[name]
    nx, ny = w["nx"], w["ny"]
    m = O.Mesh.rectangle(nx, ny)
    s = O.SmFluid(m, acc=w["acc"], flux=O.HLLC, limiter=w["limiter"], cfl=0.5)
    amr = w["kind"] == "amr"
    if amr:
        def centres():
            c = m.cells()
            return c["cx"], c["cy"]
        adapt_to_initial_data(s, name, seed, centres, s.set_state,
                              lambda crit, thr: s.set_flags(crit, thr, threads),
                              lambda ml, slope: s.refine(ml, slope), w["max_level"], w["threshold"])
    else:
        cx, cy = rect_centres(nx, ny)
        s.set_state(*initial_state(name, cx, cy, seed))
    s.update(threads)   # untimed warm-up step
    steps, cells, t0 = 0, 0, time.perf_counter()
    while True:
        if amr and steps % w["every"] == 0:
            s.set_flags(1, w["threshold"], threads)
            s.refine(w["max_level"], True)
        cells += m.n
        s.update(threads)
        steps += 1
        el = time.perf_counter() - t0
        if steps >= 2 and el >= budget_s:
            break
    return cells / el, steps, el


def cpu_baseline(name, seed, budget_s):
    """The oracle on all the host's cores (the reported baseline) and on one
    thread, each for about budget_s / 2 seconds from the same initial data."""
    w = WORKLOADS[name]
    amr = w["kind"] == "amr"
    nt = cpu_threads()
    v, steps, el = cpu_run(name, seed, budget_s / 2, nt)
    v1, steps1, el1 = cpu_run(name, seed, budget_s / 2, 1)
    return {"value": v, "unit": "cell-updates/s", "cores": nt, "kind": "port",
            "sample": f"oracle (C++ -O3 -ffp-contract=off, OpenMP {nt} threads, OMP_PROC_BIND "
                      f"{os.environ.get('OMP_PROC_BIND', 'unset')}) on the same {name} workload from the same "
                      f"initial data, {steps} steps{' (adaptation included)' if amr else ''} in {el:.2f} s",
            "single_thread": {"value": v1, "steps": steps1, "seconds": el1},
            "host": host_info()}


class ChunkProfile:
    """Per-kernel HIP-event time and algorithmic bytes accumulated over
    chunks of steps run on one mesh (n cells, phi = faces / cells)."""

    def __init__(self, ctx):
        self.ctx = ctx
        self.prev = {}
        self.ms, self.launches, self.bytes, self.cells = {}, {}, {}, {}

    def close_chunk(self, n, phi):
        cur = self.ctx.profile_read()   # cumulative since profile_enable
        kb = kernel_bytes_per_cell(phi)
        for k, (ms, cnt) in cur.items():
            pms, pcnt = self.prev.get(k, (0.0, 0))
            dms, dcnt = ms - pms, cnt - pcnt
            if dcnt <= 0:
                continue
            self.ms[k] = self.ms.get(k, 0.0) + dms
            self.launches[k] = self.launches.get(k, 0) + dcnt
            self.bytes[k] = self.bytes.get(k, 0.0) + kb.get(k, 0.0) * n * dcnt
            self.cells[k] = self.cells.get(k, 0) + n * dcnt
        self.prev = cur


def run_steps(ctx, w, steps, step0, prof=None):
    """`steps` steps from global step index step0; AMR workloads adapt
    (flags + refine) before every step whose index is a multiple of
    w["every"].  Returns the number of cell-updates."""
    cells = 0
    i = 0
    while i < steps:
        k = steps - i
        if w["kind"] == "amr":
            g = step0 + i
            if g % w["every"] == 0:
                ctx.set_flags(1, w["threshold"])
                ctx.refine(w["max_level"], slope=True)   # synchronises
                if prof is not None:
                    prof.close_chunk(prof.n, prof.phi)
                    prof.n, prof.phi = ctx.n, ctx.n_faces() / ctx.n
            k = min(k, w["every"] - (g % w["every"]))
        ctx.step(k)
        cells += ctx.n * k
        i += k
    return cells


def adjoint_run(args, w, Z):
    """Forward (with checkpoints) + backward of the discrete adjoint; FD check."""
    from zephyr_amd import problems
    nx, ny = w["nx"], w["ny"]
    ctx = Z.Context()
    ctx.mesh_rectangle(nx, ny)
    ctx.configure(acc=2, flux="hllc", limiter=w["limiter"], cfl=0.5, gamma=1.4)
    rho0, u0, v0, p0, e0 = problems.smooth_flow_field(nx, ny, args.seed)

    def state(rho):
        return rho, u0, v0, p0, problems.energy_rP(rho, p0, 1.4)

    steps = args.steps
    ctx.set_state(*state(rho0))
    ctx.step(2)                       # warm-up of the step kernels
    ctx.set_state(*state(rho0))
    ctx.adjoint_forward(4, 2, w["window"])
    ctx.adjoint_backward()            # reverse buffers allocated, adjoint kernels loaded
    ctx.set_state(*state(rho0))
    ctx.synchronize()
    t0 = time.perf_counter()
    ctx.step(steps)
    ctx.synchronize()
    t_plain = time.perf_counter() - t0
    ctx.set_state(*state(rho0))
    ctx.synchronize()
    t0 = time.perf_counter()
    J = ctx.adjoint_forward(steps, w["every"], w["window"])
    t_fwd = time.perf_counter() - t0
    t0 = time.perf_counter()
    lr, lu, lv, lp, le = ctx.adjoint_backward()
    t_bwd = time.perf_counter() - t0
    grad = lr - le * problems.energy_rP(rho0, p0, 1.4) / rho0
    # central differences of the same forward map on seeded smooth directions,
    # at the configured eps and at eps / 10 (fewer limiter / wave-speed
    # switches crossed inside +-eps)
    n_dirs = w["fd_dirs"] if not args.fd_dirs else args.fd_dirs
    fds = {e: [] for e in w["fd_eps"]}
    ads = []
    for k in range(n_dirs):
        d = problems.gaussian_random_field(nx, ny, 10_000 + k, corr_len=0.05)
        ads.append(float(np.dot(grad, d)))
        for eps in w["fd_eps"]:
            ctx.set_state(*state(rho0 + eps * d))
            Jp = ctx.adjoint_forward(steps, steps, w["window"])
            ctx.set_state(*state(rho0 - eps * d))
            Jm = ctx.adjoint_forward(steps, steps, w["window"])
            fds[eps].append((Jp - Jm) / (2 * eps))
    ads = np.array(ads)
    scale = float(np.sqrt(np.mean(ads ** 2))) if n_dirs else 1.0
    rel = {e: [abs(f - a) / max(abs(f), abs(a), 1e-300) for f, a in zip(v, ads)] for e, v in fds.items()}
    nrm = {e: [abs(f - a) / scale for f, a in zip(v, ads)] for e, v in fds.items()}
    n = nx * ny
    out = {
        "metric": "FP64 cell-updates/s (discrete adjoint of HLLC+MUSCL, reverse sweep incl. recompute)",
        "value": n * steps / t_bwd, "unit": "cell-updates/s", "n_gpus": 1, "steps": steps, "warmup": 2,
        "ms_per_step": t_bwd * 1e3 / steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": args.workload, "description": w["desc"], "cells": n, "J": J,
                   "forward_plain_ms_per_step": t_plain * 1e3 / steps,
                   "forward_with_checkpoints_ms_per_step": t_fwd * 1e3 / steps,
                   "backward_ms_per_step": t_bwd * 1e3 / steps,
                   "adjoint_over_forward_step": t_bwd / t_plain,
                   "fd_directions": n_dirs,
                   "fd_rel_err": {f"{e:g}": {"max": max(r), "median": float(np.median(r)),
                                             "within_1e-5": int(sum(x <= 1e-5 for x in r))}
                                  for e, r in rel.items() if r},
                   "fd_err_over_rms_derivative": {f"{e:g}": {"max": max(r), "median": float(np.median(r))}
                                                  for e, r in nrm.items() if r}},
    }
    print(json.dumps(out), flush=True)


# Algorithmic bytes per cell of the multi-material first-order update with
# n_mat materials: own state read and write (7 + 2 n_mat doubles each),
# code, volume, and per face slot adjacency, area and normal.
def mm_bytes_per_cell(n_mat, faces_per_cell=4):
    state = 8 * (7 + 2 * n_mat)
    return 2 * state + 4 + 8 + faces_per_cell * (4 + 8 + 16)


# ... second order, per cell: the gradient (own components in, both
# gradient rows out, adjacency and LSM weights of four slots), the
# predictor (own state and gradients in, face centres and areas, the half
# state out) and the corrector (half state and gradients of the cell, the
# old state, face geometry, the new state out); neighbours through the cache.
def mm_grad_bytes_per_cell(n_mat, faces_per_cell=4):
    k = 6 + 2 * n_mat
    return 8 * k + 2 * 8 * k + 4 + 16 + faces_per_cell * (4 + 8)


def mm_stage1_bytes_per_cell(n_mat, faces_per_cell=4):
    state, k = 8 * (7 + 2 * n_mat), 6 + 2 * n_mat
    return 2 * state + 2 * 8 * k + 4 + 24 + faces_per_cell * 24


def mm_flux2_bytes_per_cell(n_mat, faces_per_cell=4):
    state, k = 8 * (7 + 2 * n_mat), 6 + 2 * n_mat
    return 3 * state + 2 * 8 * k + 4 + 24 + faces_per_cell * (4 + 24)


# ... per composite-Riemann face (cell side): both states, code, adjacency,
# area, normal, both volumes, the master material's interface normal of one
# cell, the list entry, and the 4 + n_mat flux components written.
def mm_crp_bytes_per_face(n_mat):
    return 2 * 8 * (7 + 2 * n_mat) + 4 + 4 + 8 + 16 + 16 + 16 + 4 + 8 * (4 + n_mat)


# ... per cell of the composite-face marking: own mass fractions, code and
# four adjacency entries (neighbours' fractions through the cache).
def mm_crp_mark_bytes_per_cell(n_mat, faces_per_cell=4):
    return 8 * n_mat + 4 + faces_per_cell * 4


# ... per cell of the interface reconstruction: own volume fractions'
# inputs (rho, beta, rhos), neighbours' through the cache, vertices for
# hx / hy, and n_mat (nx, ny, p) written.
def mm_interface_bytes_per_cell(n_mat, faces_per_cell=4):
    return 8 * (1 + 2 * n_mat) + 4 + faces_per_cell * (4 + 8 + 16) + 8 * 8 + 3 * 8 * n_mat


def mm_adapt_initial(set_state, centres, set_flags, refine, seed, max_level):
    """Initial adaptation: max_level + 1 rounds of (the shock-bubble data at
    the cell centres, sub-sampled over each cell's width, flags, refine),
    then the data on the final mesh."""
    from zephyr_amd import problems
    for _ in range(max_level + 1):
        cx, cy, h = centres()
        set_state(*problems.mm_bubble_shock(cx, cy, h, seed))
        set_flags()
        refine(max_level)
    cx, cy, h = centres()
    set_state(*problems.mm_bubble_shock(cx, cy, h, seed))


def mm_steps(ctx, w, steps, step0):
    """`steps` multi-material steps from global step index step0, adapting
    before every step whose index is a multiple of w["every"] (AMR
    workloads).  Returns the number of cell-updates."""
    cells, i = 0, 0
    every = w.get("every")
    while i < steps:
        k = steps - i
        if every:
            g = step0 + i
            if g % every == 0:
                ctx.mm_set_flags()
                ctx.mm_refine(w["max_level"])   # synchronises
            k = min(k, every - (g % every))
        ctx.mm_step(k)
        cells += ctx.n * k
        i += k
    return cells


def mm_run(args, w, Z):
    """Multi-material first-order step (MmFluid with composite Riemann
    fluxes) on the shock-bubble, with or without adaptation."""
    from zephyr_amd import problems
    nx, ny = w["nx"], w["ny"]
    ctx = Z.Context()
    ctx.mesh_rectangle(nx, ny)
    mats = [problems.WATER, problems.AIR]
    ctx.mm_configure(mats, w["cfl"])
    acc = w.get("acc", 1)
    ctx.mm_set_accuracy(acc)
    ctx.mm_set_overlap(args.mm_overlap)
    amr = "max_level" in w
    st = None
    if amr:
        def centres():
            c, _, _ = ctx.mesh_download()
            return c["cx"], c["cy"], np.sqrt(c["volume"])
        mm_adapt_initial(ctx.mm_set_state, centres, ctx.mm_set_flags, ctx.mm_refine, args.seed, w["max_level"])
    else:
        cx, cy = rect_centres(nx, ny)
        st = problems.mm_bubble_shock(cx, cy, 1.0 / nx, args.seed)
        ctx.mm_set_state(*st)
    mm_steps(ctx, w, args.warmup, 0)
    ctx.synchronize()
    t0 = time.perf_counter()
    cells = mm_steps(ctx, w, args.steps, args.warmup)
    ctx.synchronize()
    el = time.perf_counter() - t0
    ctx.profile_enable(True)
    n = mm_steps(ctx, w, args.prof_steps, args.warmup + args.steps) / max(args.prof_steps, 1)
    ctx.synchronize()
    prof = ctx.profile_read()
    ctx.profile_enable(False)
    n_crp = ctx.mm_crp_faces()
    z = ctx.mm_state()
    mixed = int(np.sum(np.sum(z["beta"] > 0.0, axis=0) > 1))
    alg_of = {"mm_flux1": mm_bytes_per_cell(2) * n, "mm_crp": mm_crp_bytes_per_face(2) * n_crp,
              "mm_crp_mark": mm_crp_mark_bytes_per_cell(2) * n, "mm_interface": mm_interface_bytes_per_cell(2) * n,
              "mm_grad": mm_grad_bytes_per_cell(2) * n, "mm_stage1": mm_stage1_bytes_per_cell(2) * n,
              "mm_flux2": mm_flux2_bytes_per_cell(2) * n}
    # the deferred cells' update (overlapped steps) has no fixed cell count:
    # it is timed but not priced
    dom = max((k for k in prof if k in alg_of), key=lambda k: prof[k][0])
    avg_ms = prof[dom][0] / prof[dom][1]
    alg = alg_of[dom]
    achieved = alg / (avg_ms * 1e-3) / 1e9
    out = {
        "metric": f"FP64 cell-updates/s (multi-material {'first' if acc == 1 else 'second'} order, HLLC + PLIC "
                  "composite fluxes + PT closure" + (", AMR)" if amr else ")"),
        "value": cells / el, "unit": "cell-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": el * 1e3 / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": args.workload, "description": w["desc"], "cells": ctx.n,
                   "mean_cells_per_step": cells / args.steps,
                   "levels": np.bincount(ctx.mesh_download()[0]["level"]).tolist(),
                   "mach": problems.mm_bubble_shock_params(args.seed), "mixed_cells_final": mixed,
                   "crp_faces_last_step": n_crp,
                   "kernel_hbm_frac": {k: alg_of[k] / (prof[k][0] / prof[k][1] * 1e-3) / 1e9 / HBM_PEAK_GBS
                                       for k in prof if k in alg_of},
                   "stream_mode": args.mm_overlap,
                   "dt_final": ctx.mm_dt(),
                   "kernel_ms_avg": {k: v[0] / v[1] for k, v in prof.items()},
                   "kernel_timing": f"HIP events, {args.prof_steps} steps following the timed region"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "algorithmic_bytes_per_launch": alg, "avg_launch_ms": avg_ms},
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = mm_cpu_baseline(w, st, mats, args.cpu_budget, args.seed)
    print(json.dumps(out), flush=True)


def mm_cpu_run(w, st, mats, budget_s, seed, threads):
    """Oracle MmFluid (C++, OpenMP over cells) on the same workload (same
    initial data, same adaptation), whole steps until the budget is spent."""
    from oracle import pyoracle as O
    O.build()
    m = O.Mesh.rectangle(w["nx"], w["ny"])
    o = O.MmFluid(m, mats, w["cfl"])
    o.set_accuracy(w.get("acc", 1))
    o.set_threads(threads)
    amr = "max_level" in w
    if amr:
        def centres():
            c = m.cells()
            return c["cx"], c["cy"], np.sqrt(c["volume"])
        mm_adapt_initial(o.set_state, centres, o.set_flags, o.refine, seed, w["max_level"])
    else:
        o.set_state(*st)
    steps, cells, t0 = 0, 0, time.perf_counter()
    while time.perf_counter() - t0 < budget_s:
        if amr and steps % w["every"] == 0:
            o.set_flags()
            o.refine(w["max_level"])
        cells += m.n
        o.update()
        steps += 1
    el = time.perf_counter() - t0
    return cells / el, steps, el


def mm_cpu_baseline(w, st, mats, budget_s, seed):
    """The MmFluid oracle on all the host's cores (the reported baseline) and
    on one thread, each for about budget_s / 2 seconds."""
    amr = "max_level" in w
    nt = cpu_threads()
    v, steps, el = mm_cpu_run(w, st, mats, budget_s / 2, seed, nt)
    v1, steps1, el1 = mm_cpu_run(w, st, mats, budget_s / 2, seed, 1)
    return {"value": v, "unit": "cell-updates/s", "cores": nt, "kind": "port",
            "sample": f"oracle MmFluid (C++ -O3 -ffp-contract=off, OpenMP {nt} threads over cells) on the same "
                      f"workload from the same initial data, {steps} steps{' (adaptation included)' if amr else ''} "
                      f"in {el:.2f} s",
            "single_thread": {"value": v1, "steps": steps1, "seconds": el1},
            "host": host_info()}


def ensemble_run(args, w, Z):
    """Batched ensemble vs the looped per-problem path, same problems."""
    sizes = ensemble_sizes(w, args.seed)
    cells_per_step = sum(s * s for s in sizes)
    # batched: one context, graphs, autotuned launches
    b = Z.Context()
    b.batch_rectangles([(s, s) for s in sizes])
    b.configure(acc=w["acc"], flux="hllc", limiter=w["limiter"], cfl=0.5, gamma=1.4)
    begin, count = b.batch_layout()
    full = [np.ones(b.n), np.zeros(b.n), np.zeros(b.n), np.ones(b.n), np.full(b.n, 2.5)]
    for p, s in enumerate(sizes):
        st = ensemble_state(s, args.seed, p)
        for k in range(5):
            full[k][begin[p]:begin[p] + count[p]] = st[k]
    b.set_state(*full)
    del full
    b.step(2)
    tune = b.autotune(reps=3)
    b.graph_enable(True)
    b.step(args.warmup)
    b.synchronize()
    t0 = time.perf_counter()
    b.step(args.steps)
    b.synchronize()
    t_batch = time.perf_counter() - t0
    # roofline of the batched corrector: one profiled step (eager launches)
    b.graph_enable(False)
    b.profile_enable(True)
    b.step(6)
    b.synchronize()
    prof = b.profile_read()
    b.profile_enable(False)
    # looped: the plain path, one context per problem
    ctxs = []
    for p, s in enumerate(sizes):
        c = Z.Context()
        c.mesh_rectangle(s, s)
        c.configure(acc=w["acc"], flux="hllc", limiter=w["limiter"], cfl=0.5, gamma=1.4)
        c.set_state(*ensemble_state(s, args.seed, p))
        ctxs.append(c)
    for c in ctxs:
        c.step(args.warmup)
    for c in ctxs:
        c.synchronize()
    t0 = time.perf_counter()
    for c in ctxs:
        c.step(args.steps)
    for c in ctxs:
        c.synchronize()
    t_loop_prob = time.perf_counter() - t0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        for c in ctxs:
            c.step(1)
    for c in ctxs:
        c.synchronize()
    t_loop_step = time.perf_counter() - t0
    # the same loop with every context on one shared stream (launches serialised)
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    shared = C.c_void_p()
    if hip.hipStreamCreate(C.byref(shared)) != 0:
        raise RuntimeError("hipStreamCreate failed")
    for c in ctxs:
        c.set_stream(shared.value)
    t0 = time.perf_counter()
    for c in ctxs:
        c.step(args.steps)
    ctxs[0].synchronize()
    t_loop_one = time.perf_counter() - t0
    for c in ctxs:
        c.set_stream(0)
    hip.hipStreamDestroy(shared)
    t_loop = min(t_loop_prob, t_loop_step)   # the faster multi-stream loop is the baseline
    value = cells_per_step * args.steps / t_batch
    looped = cells_per_step * args.steps / t_loop
    phi = 2.0 + sum(2.0 * s for s in sizes) / cells_per_step   # unique faces / cells, s x s meshes
    kb = kernel_bytes_per_cell(phi)
    dom = max((k for k in ("grad_pred", "corr", "tile") if k in prof), key=lambda k: prof[k][0])
    dom_ms = prof[dom][0] / prof[dom][1]
    alg = kb[dom] * cells_per_step
    achieved = alg / (dom_ms * 1e-3) / 1e9
    fused = None
    if dom == "tile":
        fb = tile_fused_bytes_per_cell(b.launch(1)[1]) * cells_per_step
        fused = {"bytes_per_launch": fb, "achieved_GBs": fb / (dom_ms * 1e-3) / 1e9,
                 "frac": fb / (dom_ms * 1e-3) / 1e9 / HBM_PEAK_GBS}
    out = {
        "metric": "FP64 cell-updates/s (HLLC+MUSCL, batched ensemble)",
        "value": value, "unit": "cell-updates/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": t_batch * 1e3 / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": args.workload, "description": w["desc"], "problems": len(sizes),
                   "cells_per_step": cells_per_step, "phi": phi,
                   "looped_value": looped, "batched_over_looped": value / looped,
                   "looped_ms_per_step": {"per_problem_runs": t_loop_prob * 1e3 / args.steps,
                                          "interleaved_steps": t_loop_step * 1e3 / args.steps,
                                          "one_shared_stream": t_loop_one * 1e3 / args.steps},
                   "batched_over_looped_one_stream": t_loop_one / t_batch,
                   "launch": {str(st): b.launch(st) for st in (1, 2)},
                   "tile_persistent": b.tile_schedule(),
                   "autotune_ms": tune.tolist(),
                   "kernel_ms_avg": {k: v[0] / v[1] for k, v in prof.items()}},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None, "algorithmic_bytes_per_launch": alg,
                     "avg_launch_ms": dom_ms, "fused_kernel_minimum": fused},
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = ensemble_cpu_baseline(w, sizes, args.seed, args.cpu_budget)
    print(json.dumps(out), flush=True)


def ensemble_cpu_run(w, sizes, seed, budget_s, threads):
    """Oracle on the ensemble's problems, in order, two steps each, until the
    budget is spent; returns (cell-updates/s, problems, cell-updates, seconds)."""
    from oracle import pyoracle as O
    cells, probs, el = 0, 0, 0.0
    for p, s in enumerate(sizes):
        m = O.Mesh.rectangle(s, s)
        f = O.SmFluid(m, acc=w["acc"], flux=O.HLLC, limiter=w["limiter"], cfl=0.5)
        f.set_state(*ensemble_state(s, seed, p))
        t0 = time.perf_counter()
        f.update(threads)
        f.update(threads)
        el += time.perf_counter() - t0
        cells += 2 * s * s
        probs += 1
        if el >= budget_s:
            break
    return cells / el, probs, cells, el


def ensemble_cpu_baseline(w, sizes, seed, budget_s):
    """The oracle on all the host's cores (the reported baseline) and on one
    thread, each for about budget_s / 2 seconds of steps."""
    nt = cpu_threads()
    v, probs, cells, el = ensemble_cpu_run(w, sizes, seed, budget_s / 2, nt)
    v1, probs1, cells1, el1 = ensemble_cpu_run(w, sizes, seed, budget_s / 2, 1)
    return {"value": v, "unit": "cell-updates/s", "cores": nt, "kind": "port",
            "sample": f"oracle (C++ -O3 -ffp-contract=off, OpenMP {nt} threads): two steps of each of the "
                      f"first {probs} ensemble problems ({cells} cell-updates, steps only) in {el:.2f} s",
            "single_thread": {"value": v1, "problems": probs1, "seconds": el1},
            "host": host_info()}


# Compulsory bytes per cell of the three-dimensional step kernels: every
# array the kernel touches, once (neighbour values served from cache).
# Slot-major mesh: per slot type + adj (8), face centre (24), area (8) and,
# for the gradient, the weight (8); cell centre 24, volume 8, incircle 8;
# state rho, u, v, w, p, e (48); gradients 3 x 5 doubles (120); wait 4.
# The first-order update and the corrector also reduce the next step's CFL
# candidate (incircle, 8), so the standalone reduction runs only after a new
# state is set.
CUBE_BYTES_SLOTS = {
    "dt": 40 + 8,
    "flux1": 48 + 6 * 16 + 8 + 48 + 8,
    "grad_pred": 48 + 24 + 6 * 48 + 4 + 8 + 120 + 48,
    "corr": 48 + 48 + 120 + 4 + 4 + 24 + 8 + 6 * 40 + 48 + 8,
}
# With the Cuboid's geometry tables (the library's default on a Cuboid) the
# mesh costs only the incircle per cell: the tables (25 (nx + ny + nz)
# doubles) stay in cache.
CUBE_BYTES = {
    "dt": 40 + 8,
    "flux1": 48 + 48 + 8,
    "grad_pred": 48 + 4 + 120 + 48,
    "corr": 48 + 48 + 120 + 4 + 4 + 48 + 8,
}


def cube_state(cx, cy, cz, seed):
    rng = np.random.default_rng(seed)
    k = rng.uniform(2.0, 6.0, size=(4, 3))
    ph = rng.uniform(0.0, 2.0 * np.pi, size=4)
    r = np.sqrt((cx - 0.5) ** 2 + (cy - 0.5) ** 2 + (cz - 0.5) ** 2)
    inside = r < 0.25
    rho = np.where(inside, 1.5, 1.0) * (1.0 + 0.01 * np.sin(k[0, 0] * cx + k[0, 1] * cy + k[0, 2] * cz + ph[0]))
    u = 0.05 * np.sin(k[1, 0] * cx + k[1, 1] * cy + k[1, 2] * cz + ph[1])
    v = 0.05 * np.sin(k[2, 0] * cx + k[2, 1] * cy + k[2, 2] * cz + ph[2])
    w = 0.05 * np.sin(k[3, 0] * cx + k[3, 1] * cy + k[3, 2] * cz + ph[3])
    p = np.where(inside, 10.0, 1.0)
    return rho, u, v, w, p, p / (0.4 * rho)


def cube_cpu_baseline(w, seed, budget_s):
    """The oracle's SmFluid3 (C++, OpenMP) on the same mesh and initial data,
    whole steps until about budget_s seconds are spent."""
    from oracle import pyoracle as O
    nt = cpu_threads()
    m = O.Mesh3(w["nx"], w["ny"], w["nz"])
    f = O.SmFluid3(m, flux=O.HLLC, acc=w["acc"], limiter=w["limiter"], cfl=0.5)
    c = m.cells()
    f.set_state(*cube_state(c["cx"], c["cy"], c["cz"], seed))
    steps, el = 0, 0.0
    while el < budget_s:
        t0 = time.perf_counter()
        f.update(nt)
        el += time.perf_counter() - t0
        steps += 1
    n = m.n
    return {"value": n * steps / el, "unit": "cell-updates/s", "cores": nt, "kind": "port",
            "sample": f"oracle SmFluid3 (C++ -O3 -ffp-contract=off, OpenMP {nt} threads): the first {steps} "
                      f"steps of the same {w['nx']}x{w['ny']}x{w['nz']} workload from the same initial data in "
                      f"{el:.2f} s", "host": host_info()}


def cube_run(args, w, Z):
    ctx = Z.Context3()
    ctx.mesh_cuboid(w["nx"], w["ny"], w["nz"])
    ctx.configure(acc=w["acc"], flux="hllc", limiter=w["limiter"], cfl=0.5, gamma=1.4)
    cells, _ = ctx.mesh()
    ctx.set_state(*cube_state(cells["cx"], cells["cy"], cells["cz"], args.seed))
    del cells
    n = ctx.n
    ctx.step(args.warmup)
    ctx.synchronize()
    t0 = time.perf_counter()
    ctx.step(args.steps)
    ctx.synchronize()
    el = time.perf_counter() - t0
    if args.copy_calibration:
        # 3 device-to-device copies of n doubles: 8n bytes read and 8n bytes
        # written per copy dispatch (the scale of the HBM counters)
        import ctypes as C
        hip = C.CDLL("libamdhip64.so")
        src, dst = C.c_void_p(), C.c_void_p()
        nbytes = C.c_size_t(8 * n)
        if hip.hipMalloc(C.byref(src), nbytes) != 0 or hip.hipMalloc(C.byref(dst), nbytes) != 0:
            raise RuntimeError("hipMalloc failed")
        for _ in range(3):
            if hip.hipMemcpy(dst, src, nbytes, 3) != 0:   # hipMemcpyDeviceToDevice
                raise RuntimeError("hipMemcpy failed")
        hip.hipDeviceSynchronize()
        hip.hipFree(src)
        hip.hipFree(dst)
    # kernel timing pass: the workload continues with events around each launch
    ctx.profile_enable(True)
    ctx.step(min(args.prof_steps, 20))
    prof = ctx.profile_read()
    ctx.profile_enable(False)
    avg = {k: v[0] / v[1] for k, v in prof.items()}
    dom = max((k for k in ("grad_pred", "corr", "flux1") if k in avg), key=lambda k: avg[k])
    tables = ctx.geometry_tables()
    cube_bytes = CUBE_BYTES if tables else CUBE_BYTES_SLOTS
    alg = cube_bytes[dom] * n
    achieved = alg / (avg[dom] * 1e-3) / 1e9
    value = n * args.steps / el
    # HBM bytes per launch from the counter runs of the same kernel sources
    traffic, traffic_detail = None, {"measured_in": f"no counter run of k3_{dom} for this workload"}
    tpath = os.path.join(HERE, "profiles", "traffic.json")
    t = json.load(open(tpath)).get(args.workload, {}).get("k3_" + dom) if os.path.exists(tpath) else None
    if t:
        traffic_detail = dict(t)
        src_sha = kernel_sources_sha("k3_" + dom)
        if t.get("kernel_sources_sha") == src_sha:
            traffic = t["bytes_per_cell"] * n
            traffic_detail["measured_in"] = "separate counter runs of this command on the same kernel sources"
        else:
            traffic_detail["measured_in"] = (f"stale: counter runs of kernel sources {t.get('kernel_sources_sha')}, "
                                             f"this run executes {src_sha}; traffic not reported")
    out = {
        "metric": "FP64 cell-updates/s (HLLC+MUSCL, 3D Cuboid)",
        "value": value, "unit": "cell-updates/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": el * 1e3 / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": args.workload, "description": w["desc"], "cells": n,
                   "kernel_ms_avg": avg, "geometry_tables": tables, "bytes_per_cell": cube_bytes,
                   "step_compulsory_GBs": sum(cube_bytes[k] for k in ("grad_pred", "corr")) * n
                   / (el / args.steps) / 1e9},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_detail": traffic_detail,
                     "algorithmic_bytes_per_launch": alg, "avg_launch_ms": avg[dom]},
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cube_cpu_baseline(w, args.seed, args.cpu_budget)
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (default: the workload's own, 500 or 200)")
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="amr256_muscl", choices=sorted(WORKLOADS))
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--cpu-budget", type=float, default=10.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--fd-dirs", type=int, default=0, help="adjoint workload: finite-difference directions")
    ap.add_argument("--prof-steps", type=int, default=100,
                    help="steps of the event-timed kernel pass after the timed region")
    ap.add_argument("--mm-overlap", type=int, default=0, choices=(0, 1, 2),
                    help="multi-material workloads: 0 one stream, 1 interface on a second stream, 2 also the "
                         "composite fluxes beside the other cells' update")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run device-to-device copies of the state (known bytes) for HBM counter scaling")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only; --gpus {args.gpus} is not supported "
              f"(the sharded design is described in DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import zephyr_amd as Z

    w = WORKLOADS[args.workload]
    if args.steps is None:
        args.steps = w.get("steps", 500)
    if w["kind"] == "ensemble":
        return ensemble_run(args, w, Z)
    if w["kind"] == "adjoint":
        return adjoint_run(args, w, Z)
    if w["kind"] == "mm":
        return mm_run(args, w, Z)
    if w["kind"] == "cube":
        return cube_run(args, w, Z)
    nx, ny, acc, limiter = w["nx"], w["ny"], w["acc"], w["limiter"]
    ctx = Z.Context()
    ctx.mesh_rectangle(nx, ny)
    ctx.configure(acc=acc, flux="hllc", limiter=limiter, cfl=0.5, gamma=1.4)
    if w["kind"] == "amr":
        def centres():
            c, _, _ = ctx.mesh_download()
            return c["cx"], c["cy"]
        adapt_to_initial_data(ctx, args.workload, args.seed, centres, ctx.set_state, ctx.set_flags,
                              lambda ml, slope: ctx.refine(ml, slope), w["max_level"], w["threshold"])
        # device capacity for the adapted meshes of the run (AmrCells::reserve_amr):
        # the timed adaptations then allocate nothing (the mesh peaks near
        # 2 x the base cell count on this workload)
        ctx.reserve(RESERVE_PER_BASE_CELL * nx * ny)
    else:
        cx, cy = rect_centres(nx, ny)
        ctx.set_state(*initial_state(args.workload, cx, cy, args.seed))

    run_steps(ctx, w, args.warmup, 0)
    ctx.synchronize()
    # launch configuration (layout, workgroup size per stage) picked by the
    # library's autotuner on the warmed-up state; the state does not advance
    tune = ctx.autotune(reps=3)
    ctx.synchronize()

    if args.copy_calibration:
        # 5 device-to-device copies of n doubles (the init layer) into fresh
        # buffers: 8n bytes read and 8n bytes written per copy dispatch.
        import ctypes as C
        hip = C.CDLL("libamdhip64.so")
        bufs = [C.c_void_p() for _ in range(5)]
        for b in bufs:
            if hip.hipMalloc(C.byref(b), C.c_size_t(8 * ctx.n)) != 0:
                raise RuntimeError("hipMalloc failed")
        for _ in range(3):
            Z._check(Z.lib().zfv_state_download_device(ctx.h, *bufs))
        ctx.synchronize()
        for b in bufs:
            hip.hipFree(b)
        print(f"copy calibration: n = {ctx.n} doubles per copy", file=sys.stderr, flush=True)

    # timed region: exactly --steps steps, no per-kernel events (an event pair
    # around every launch adds an idle gap between dependent kernels)
    ctx.synchronize()
    t0 = time.perf_counter()
    cells = run_steps(ctx, w, args.steps, args.warmup)
    ctx.synchronize()
    el = time.perf_counter() - t0

    # kernel timing pass: the workload continues for --prof-steps steps with
    # HIP events around every step-kernel launch (on the context's stream)
    prof = ChunkProfile(ctx)
    prof.n, prof.phi = ctx.n, ctx.n_faces() / ctx.n
    ctx.profile_enable(True)
    ctx.synchronize()
    run_steps(ctx, w, args.prof_steps, args.warmup + args.steps, prof)
    ctx.synchronize()
    prof.close_chunk(prof.n, prof.phi)
    ctx.profile_enable(False)

    value = cells / el
    step_kernels = ("flux1",) if acc == 1 else ("grad_pred", "corr", "tile")
    dom = max(step_kernels, key=lambda k: prof.ms.get(k, 0.0))
    dom_ms = prof.ms[dom] / prof.launches[dom]
    alg_bytes = prof.bytes[dom] / prof.launches[dom]
    achieved = alg_bytes / (dom_ms * 1e-3) / 1e9
    # the tiled step fuses the gradient, predictor and corrector: the
    # stage-structured bytes above are the workload's accounting; beside them
    # the bytes the fused kernel itself must move
    fused = None
    if dom == "tile":
        fb = tile_fused_bytes_per_cell(ctx.launch(1)[1]) * prof.cells[dom] / prof.launches[dom]
        fused = {"bytes_per_launch": fb, "achieved_GBs": fb / (dom_ms * 1e-3) / 1e9,
                 "frac": fb / (dom_ms * 1e-3) / 1e9 / HBM_PEAK_GBS}
    step_bytes = sum(prof.bytes.get(k, 0.0) for k in step_kernels)
    step_ms = sum(prof.ms.get(k, 0.0) for k in step_kernels)

    # HBM bytes per launch of the dominant kernel.  rocprofv3 counters cannot
    # be read from inside this run; they come from separate FETCH_SIZE /
    # WRITE_SIZE counter runs of this same command summarised in
    # profiles/traffic.json (KiB per dispatch, scaled by the copy-kernel
    # calibration; per cell for the adaptive workload, whose launches vary
    # in size), and are used only when those runs measured the kernel
    # sources this run executes (same source hash); otherwise null.
    traffic, traffic_detail = None, None
    tpath = os.path.join(HERE, "profiles", "traffic.json")
    src_sha = kernel_sources_sha(dom)
    # the entry of the launch form in use: kernel, layout, workgroup size
    # (or tile shape), whose register use and access pattern set the bytes
    lay, blk = ctx.launch(2 if dom == "corr" else 1)
    tkey = f"{dom}:{lay}:{blk}"
    if os.path.exists(tpath):
        t = json.load(open(tpath)).get(args.workload, {}).get(tkey)
        if t:
            traffic_detail = dict(t)
            if t.get("kernel_sources_sha") == src_sha:
                per_launch_cells = prof.cells[dom] / prof.launches[dom]
                traffic = t["bytes_per_cell"] * per_launch_cells
                traffic_detail["measured_in"] = "separate counter runs of this command on the same kernel sources"
            else:
                traffic_detail["measured_in"] = (f"stale: counter runs of kernel sources {t.get('kernel_sources_sha')}, "
                                                 f"this run executes {src_sha}; traffic not reported")
        else:
            traffic_detail = {"measured_in": f"no counter run of {tkey} (kernel:layout:workgroup) for this workload"}

    scheme = "MUSCL+HLLC" if acc == 2 else "first-order HLLC"
    out = {
        "metric": f"FP64 cell-updates/s (HLLC+MUSCL{', AMR' if w['kind'] == 'amr' else ''})"
                  if acc == 2 else "FP64 cell-updates/s (HLLC first order)",
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
        "config": {"workload": args.workload, "description": w["desc"], "scheme": scheme,
                   "limiter": limiter, "cells_final": ctx.n, "mean_cells_per_step": cells / args.steps,
                   "phi_final": ctx.n_faces() / ctx.n,
                   "kernel_ms_avg": {k: prof.ms[k] / prof.launches[k] for k in prof.ms},
                   "kernel_launches": dict(prof.launches),
                   "kernel_timing": f"HIP events, {args.prof_steps} steps following the timed region",
                   "launch": {str(st): ctx.launch(st) for st in ((0,) if acc == 1 else (1, 2))},
                   "tile_persistent": ctx.tile_schedule(),
                   "autotune_ms": tune.tolist(),
                   "step_kernels_hbm_frac": step_bytes / (step_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                   "step_kernels_ms_share_of_wall": (step_ms / max(args.prof_steps, 1)) / (el * 1e3 / args.steps)},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "algorithmic_bytes_per_launch": alg_bytes, "avg_launch_ms": dom_ms,
                     "fused_kernel_minimum": fused, "traffic_detail": traffic_detail},
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.workload, args.seed, args.cpu_budget)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
