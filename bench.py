"""Ladder-cycle throughput of the HIP stepper on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload open_3band_256x128x32]

A "step" of this benchmark is one ladder cycle: lMax = 40 time steps of the release-2.0 step
(the n that TDGL's ladder hands to its stepper hook, gtdgl/src/protocol.jl:28, 409), with the
state resident on the device.  The default workload is the largest one-GPU configuration: three
anisotropic bands on 256x128x32 cells with an 8-cell insulating buffer, the open boundary and the
PCG field solve with the Float32 preconditioner (juliatdgl_amd.configs.open_3band_256x128x32); the
two-band KWT run on 128x64x16 cells is measured beside it (config.secondary).
Prints one JSON line: cell-cycles/s (nx·ny·nz × cycles / s), the roofline of the dominant
kernel (HIP events around its launches over the timed cycles), and the CPU oracle timed on a
bounded sample of the same workload on this host, with one thread and with every host thread.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

from juliatdgl_amd import configs  # noqa: E402
from juliatdgl_amd.hip import HipStepper, device_count  # noqa: E402
from juliatdgl_amd.state import seeded_state  # noqa: E402

HBM_PEAK = 8.0e12       # B/s, MI355X HBM3E
FP64_PEAK = 78.6e12     # flop/s, MI355X FP64 (vector and matrix)
FP32_MFMA_PEAK = 157.3e12  # flop/s, MI355X dense FP32 matrix
LMAX = 40


def kernel_bytes(st, name: str, launches: int, nsteps: int, stats: dict) -> float:
    """Algorithmic bytes (each array a kernel must touch counted once) summed over the launches
    of one kernel class in `nsteps` time steps."""
    gr = st.grid
    N = st.N
    refine = getattr(st, "pcg_refine", False)
    interior = (gr.x2 - gr.x1) * (gr.y2 - gr.y1) * (gr.z2 - gr.z1)
    nodes = (gr.Nx + 1) * (gr.Ny + 1) * (gr.Nz + 1)
    rng = (gr.Nx - 2) * (gr.Ny - 1) * (gr.Nz - 1) + (gr.Nx - 1) * (gr.Ny - 2) * (gr.Nz - 1) + \
        (gr.Nx - 1) * (gr.Ny - 1) * (gr.Nz - 2)
    if name == "band_stage":
        # S, D, the other bands, f, three links (+ Y_{j-2} and Ω on most stages)
        per = 16 + 16 + 16 * (N - 1) + 8 + 48 + 16 + 8
        return per * interior * launches
    if name == "sep_transform_f64":
        return 16.0 * rng * launches
    if name == "sep_transform_f32":
        if stats.get("folded"):
            # the quarter transforms: Float32 quarters in and out, 8 B per quarter element
            return 8.0 * quarter_total(gr) * launches
        # per PCG solve of `it` iterations: it + 1 forward launches (FP64 in, FP32 out: 12 B per
        # element), the first inverse launch (FP32 in, FP64 out, r read for r·z: 20 B) and it
        # inverse launches of the flexible iteration (r and z_old read: 28 B)
        it = stats.get("iters_per_solve", 0.0)
        per = (12.0 * (it + 1) + 20.0 + 28.0 * it) / (2.0 * (it + 1))
        return per * rng * launches
    if name == "sep_fold":
        if refine:
            # the inner solves' unfolds: Float32 quarters, r, z_old in, z out (16 B)
            return 16.0 * rng * launches
        if stats.get("folded"):
            # per solve of `it` iterations: one fold (FP64 in, FP32 quarters out: 12 B, the start's;
            # the iterations' folds ride on the x/r update), the start's unfold (FP32 in, FP64 z
            # out, r read for r·z: 20 B) and `it` unfolds that also read z_old (28 B)
            it = stats.get("iters_per_solve", 0.0)
            return (12.0 + 20.0 + 28.0 * it) / (it + 2.0) * rng * launches
        # the fold (FP64 in, FP32 quarters out: 12 B) and the unfold (FP32 in, FP64 z out, r and
        # z_old read for the dots: 28 B), one each per preconditioner
        return 20.0 * rng * launches
    if name == "sep_modes":
        return 8.0 * (quarter_total(gr) if stats.get("folded") else rng) * launches
    if name == "pcg_apply_A":
        # z, P_old, D read, P, Q written (and z_old = z written unless the folded solve alternates
        # the two buffers); the refinement's inner iterations in Float32
        return (20.0 if refine else 40.0 if stats.get("folded") else 48.0) * rng * launches
    if name == "pcg_vector":
        # x, P, r, Q read, x, r written (folded: and the Float32 quarters of r's fold written)
        return (28.0 if refine else 52.0 if stats.get("folded") else 48.0) * rng * launches
    if name == "current_field":
        return (16 * N + 48 + 24 * 6 + 24 + 24) * nodes * launches
    return 0.0


def ranges(gr):
    return [(gr.Nx - 2, gr.Ny - 1, gr.Nz - 1), (gr.Nx - 1, gr.Ny - 2, gr.Nz - 1), (gr.Nx - 1, gr.Ny - 1, gr.Nz - 2)]


def quarters(gr):
    """The parity quarters of the folded transforms: per component (ceil(Lx/2), ceil(Ly/2), Lz)."""
    return [((ex[0] + 1) // 2, (ex[1] + 1) // 2, ex[2]) for ex in ranges(gr)]


def quarter_total(gr):
    return sum(4 * h0 * h1 * e2 for h0, h1, e2 in quarters(gr))


def kernel_flops(st, name: str, launches: int, folded: bool = False) -> float:
    """Algorithmic flops of the launches of one kernel class: a fused x-y transform launch of the
    separable solve multiplies every element by an (Lx x Lx) then an (Ly x Ly) matrix, 2 (Lx + Ly)
    flop per element; folded, each of the four quarters of a component by its half-size bases,
    2 (ceil(Lx/2) + ceil(Ly/2)) flop per quarter element (half the work).  The folded Float32
    preconditioner's z pass takes every quarter element by an (Lz x Lz) matrix forward and back,
    4 Lz flop per element."""
    if name == "sep_modes" and folded and st.field_mode == 2:
        return launches * sum(4 * h0 * h1 * e2 * 4.0 * e2 for h0, h1, e2 in quarters(st.grid))
    if name == "sep_transform_f32" and folded:
        return launches * sum(4 * 2.0 * h0 * h1 * e2 * (h0 + h1) for h0, h1, e2 in quarters(st.grid))
    if name.startswith("sep_transform"):
        per = sum(2.0 * ex[0] * ex[1] * ex[2] * (ex[0] + ex[1]) for ex in ranges(st.grid))
        return per * launches
    return 0.0


def cycle_bytes_per_node_step(st, iters_per_step: float) -> float:
    """B_alg per padded node and step of the whole step (every array a fused phase touches counted
    once): the bands' stages, supercurrent + field increment, the PCG (per iteration with the
    Float32 preconditioner) and links + curl, as tabulated for this path."""
    s = sum(max(x, 1) for x in st.rkl)
    N = st.N
    band = 2 * (16 * (N + 1) + 8 + 72) + 152 * (s - 2)
    field = 16 * N + 192
    if st.field_mode == 2 and getattr(st, "pcg_refine", False):
        # Float32 inner iterations (their vectors, operator and D at half the bytes of the FP64
        # ones: 176 for 288), and about three FP64 rounds a step (residual, check, x += d)
        pcg = 300 + 176 * iters_per_step + 3 * 150
    else:
        pcg = 300 + 288 * iters_per_step if st.field_mode == 2 else (168 if st.field_mode == 1 else 0)
    return band + field + pcg + 96


def measured_traffic(kernel: str, workload: str):
    """HBM bytes per launch of `kernel` on `workload` from the committed rocprofv3 PMC summary
    (FETCH_SIZE and WRITE_SIZE per dispatch, each scaled by what the counter reads for a copy
    of known size), or None where there is none."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload}_traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def copy_probe(mib: int) -> None:
    """One device-to-device copy of `mib` MiB through the HIP runtime (its copy kernel): the
    reference dispatch whose FETCH_SIZE and WRITE_SIZE counters scale the counted HBM traffic."""
    import ctypes
    try:
        hip = ctypes.CDLL("libamdhip64.so")
    except OSError:
        hip = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so")
    n = ctypes.c_size_t(mib << 20)
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    for call in (lambda: hip.hipMalloc(ctypes.byref(a), n), lambda: hip.hipMalloc(ctypes.byref(b), n),
                 lambda: hip.hipMemset(a, 1, n), lambda: hip.hipDeviceSynchronize(),
                 lambda: hip.hipMemcpy(b, a, n, 3), lambda: hip.hipDeviceSynchronize(),  # 3: device to device
                 lambda: hip.hipFree(a), lambda: hip.hipFree(b)):
        rc = call()
        if rc != 0:
            raise RuntimeError(f"copy probe: HIP error {rc}")
    print(f"[bench] copy probe: {mib} MiB read + {mib} MiB written", file=sys.stderr)


def bench_sweep(args) -> int:
    """The batched current sweep: 64 independent plain FP64 runs (32 currents in [0.05, 0.4] x 2
    polarities, grids uniform on multiples of 4 in [32,160]x[16,96]x[4,24], notch depth in [0.1, 0.4]
    of the width, per-run seeds from one master seed), each `steps` ladder cycles measured on the
    device.  Runs/hour one after the other and `workers` at a time, whether every run's record is
    the same alone and in the mix, and p50/p99 latency of the same runs served from a Poisson stream
    (continuous admission by HBM footprint)."""
    from juliatdgl_amd import sweep
    specs = sweep.sweep_specs(32, master=2026)
    work = sweep.Cycles(cycles=args.steps)
    cells = sum(s.nx * s.ny * s.nz for s in specs)
    print(f"[bench] sweep64: {len(specs)} runs, {cells} cells, {args.steps} cycles each", file=sys.stderr, flush=True)
    sweep.run_batch(specs[:2], sweep.Cycles(cycles=1), workers=2)  # the device and the library warm
    seq, t_seq = sweep.run_batch(specs, work, workers=1)
    print(f"[bench] one at a time: {t_seq:.2f} s", file=sys.stderr, flush=True)
    bat, t_bat = sweep.run_batch(specs, work, workers=args.workers)
    print(f"[bench] {args.workers} at a time: {t_bat:.2f} s", file=sys.stderr, flush=True)
    kb, t_kb = sweep.run_batched(specs, work, group=len(specs), workers=args.workers)
    print(f"[bench] batched kernels (one launch per kernel over all runs of a solve path): {t_kb:.2f} s", file=sys.stderr,
          flush=True)
    kb16, t_kb16 = sweep.run_batched(specs, work, group=16, workers=args.workers)
    print(f"[bench] batched kernels, groups of 16 set up while the last runs: {t_kb16:.2f} s", file=sys.stderr, flush=True)
    kbl, t_kbl = sweep.run_batched(specs, work, group=16, workers=args.workers, lanes=4)
    print(f"[bench] batched kernels, four groups of 16 at once: {t_kbl:.2f} s", file=sys.stderr, flush=True)
    kbl8, t_kbl8 = sweep.run_batched(specs, work, group=8, workers=args.workers, lanes=4)
    print(f"[bench] batched kernels, groups of 8, four at once: {t_kbl8:.2f} s", file=sys.stderr, flush=True)
    same = all(seq[q] == bat[q] == kb[q] == kb16[q] == kbl[q] == kbl8[q] for q in seq)
    rate = len(specs) / t_bat
    srv = sweep.Server(budget_bytes=200e9, slots=args.workers)
    for s, t in zip(specs, sweep.poisson_arrivals(len(specs), rate=0.9 * rate, seed=5)):
        srv.submit(s, t)
    t0 = time.perf_counter()
    done = srv.run(work)
    t_srv = time.perf_counter() - t0
    lat = sweep.percentiles(d["latency"] for d in done.values())
    same_srv = all(done[q]["result"] == seq[q] for q in seq)
    print(f"[bench] served: {t_srv:.2f} s, p50 {lat['p50']:.3f} s, p99 {lat['p99']:.3f} s", file=sys.stderr, flush=True)
    line = {"metric": "batched-sweep runs/hr", "value": 3600.0 * len(specs) / t_bat, "unit": "runs/hr", "n_gpus": 1,
            "steps": args.steps, "warmup": 1, "ms_per_step": 1e3 * t_bat / args.steps, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded psi per run)",
            "config": {"workload": "64-run current sweep: plain FP64 two-band runs (kappa 2.5, separable solve, "
                                   "bar field, dt 0.01), grids 32..160 x 16..96 x 4..24, notch depth 0.1..0.4, "
                                   f"J in +-[0.05, 0.4], {args.steps} ladder cycles of 40 steps per run",
                       "runs": len(specs), "cells": cells, "workers": args.workers,
                       "sequential_runs_per_hr": 3600.0 * len(specs) / t_seq, "speedup": t_seq / t_bat,
                       "batched_kernels_runs_per_hr": 3600.0 * len(specs) / t_kb,
                       "batched_kernels_groups16_runs_per_hr": 3600.0 * len(specs) / t_kb16,
                       "batched_kernels_4x16_at_once_runs_per_hr": 3600.0 * len(specs) / t_kbl,
                       "batched_kernels_4x8_at_once_runs_per_hr": 3600.0 * len(specs) / t_kbl8,
                       "cell_cycles_per_s_batched": cells * args.steps / t_bat,
                       "same_alone_and_in_the_mix": bool(same and same_srv),
                       "served_p50_latency_s": lat["p50"], "served_p99_latency_s": lat["p99"],
                       "served_arrival_rate_per_s": 0.9 * rate, "served_runs_per_hr": 3600.0 * len(specs) / t_srv},
            "roofline": None, "cpu_baseline": None}
    print(json.dumps(line), flush=True)
    return 0


REFINE = {"pcg_inner_tol": 1e-4, "pcg_rounds": 3}  # the refinement workload's inner tolerance and rounds

WORKLOADS = {
    "open_3band_256x128x32": (
        lambda: configs.open_3band_256x128x32(),
        "three-band anisotropic (m_z = 4 m_xy) TDGL, 256x128x32 cells, dx=0.5, buf 8 insulating (sigma_vac/sigma_n=0.01), "
        "PCG tol 1e-8 (Float32 preconditioner, FP64 outer), open boundary (FFT convolution, every t_GL), notch, H=0.2, "
        "J=0.1, ladder cycle = 40 steps", "notch mask"),
    "open_3band_256x128x32_refine": (
        lambda: configs.open_3band_256x128x32(pcg_refine=True, **REFINE),
        "three-band anisotropic (m_z = 4 m_xy) TDGL, 256x128x32 cells, dx=0.5, buf 8 insulating (sigma_vac/sigma_n=0.01), "
        "PCG tol 1e-8 by FP64 iterative refinement around Float32 inner PCG solves (each to 1e-4 of its starting "
        "residual by default, Float32 preconditioner), open boundary (FFT convolution, every t_GL), notch, H=0.2, J=0.1, ladder "
        "cycle = 40 steps", "notch mask"),
    "kwt_pcg_128x64x16": (
        lambda: configs.kwt_pcg_128x64x16(),
        "two-band KWT (u=Gamma/sigma_n=5.79, gkwt=10) TDGL, 128x64x16 cells, dx=0.5, buf 2 insulating "
        "(sigma_vac/sigma_n=0.01), PCG tol 1e-10 (Float32 preconditioner, FP64 outer), 30-degree wedge, H=0.2, J=0.1, "
        "dt_scale=0.1, ladder cycle = 40 steps", "30-degree sloped wedge mask"),
    "pcg_wedge_96x48x24": (
        lambda: configs.pcg_wedge_96x48x24(),
        "the reference's validation run file pcg96.toml: mgb2_model_II at kappa=2, 96x48x24 cells, dx=0.5, buf 2 "
        "insulating (sigma_vac=0.01), PCG tol 1e-8 (Float32 preconditioner, FP64 outer), WEDGE, J=0.3, H=0.03, ladder "
        "cycle = 40 steps", "catalogue WEDGE mask"),
    "plain_wedge_96x48x24": (
        lambda: configs.plain_wedge_96x48x24(),
        "the reference's validation run file plain96.toml: mgb2_model_II at kappa=2, 96x48x24 cells, dx=0.5, buf 1, "
        "separable implicit field solve, WEDGE, J=0.3, H=0.03, ladder cycle = 40 steps", "catalogue WEDGE mask"),
}
# the reference's published T4 figures for its validation run files (ms per time step, median of three batches of
# 20 steps; gpu/TDGLGPU/validation/README.md:41-71)
PUBLISHED = {"pcg_wedge_96x48x24": {"t4_batch_ms_per_step": 26.903, "t4_cycle_ms_per_step": 21.040,
                                    "laptop_1core_ms_per_step": 146.370},
             "plain_wedge_96x48x24": {"t4_batch_ms_per_step": 7.295, "t4_cycle_ms_per_step": 4.816,
                                      "laptop_1core_ms_per_step": 22.574}}


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_threads() -> int:
    """The host threads this process may use (its affinity set), capped at 16 (the box's CPU share)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_sample(name, st, dev, H, HJ, nsteps):
    """The oracle's legs of the CPU baseline: one thread, then every host thread.  On the workloads
    whose whole grid the oracle steps in seconds, `nsteps` steps from the device's state and PCG
    history after the warm-up.  On the 256x128x32 workload one oracle step takes minutes (its dense
    separable transforms run on the CPU's vector units, and its open-boundary convolution is a
    direct sum), so the sample is the same workload on 64x32x8 cells: two untimed steps (the cold
    PCG start and the first open-boundary recomputation, which recurs every 39 steps), then
    `nsteps` timed ones, scaled by the sample's own cells."""
    from oracle.oracle import Oracle, set_threads
    if name == "open_3band_256x128x32":
        st2, p2 = configs.open_3band_256x128x32(nx=64, ny=32, nz=8)
        H2, HJ2 = p2["H"], st2.selffield(p2["J"])
        s0 = seeded_state(st2.model, st2.grid, st2.masks, H2, seed=20261202)
        pre = Oracle(st2)
        pre.set_state(s0)
        set_threads(host_threads())
        pre.step(2, H2, HJ2)
        start, pcg, opn = pre.get_state(), pre.get_pcg(), pre.open_state()
        sample = (f"the same workload on 64x32x8 cells (buf 8: {st2.grid.Nx}x{st2.grid.Ny}x{st2.grid.Nz}), {nsteps} time "
                  f"steps after two untimed ones, scaled by its {st2.grid.cells} cells; the open boundary's recomputation "
                  f"(every 39 steps) is outside the timed steps")
        stx, Hx, HJx, cells = st2, H2, HJ2, st2.grid.cells
    else:
        start = dev.download()
        pcg = dev.get_pcg() if st.field_mode == 2 else None
        opn = dev.get_open() if st.selffield_bc == "open" else None
        sample = f"{nsteps} time steps (of 40 per cycle) of the same workload from the GPU's state after the warm-up"
        stx, Hx, HJx, cells = st, H, HJ, st.grid.cells
    legs = []
    for threads in (1, host_threads()):
        set_threads(threads)
        o = Oracle(stx)
        o.set_state(start)
        if pcg is not None:
            o.set_pcg(pcg)
        if opn is not None:
            o.set_open(*opn)
        t = time.perf_counter()
        o.step(nsteps, Hx, HJx)
        dt = time.perf_counter() - t
        legs.append((threads, cells * (nsteps / LMAX) / dt, dt))
        print(f"[bench] cpu oracle {threads} thread(s): {legs[-1][1]:.3e} cell-cycles/s ({dt:.1f}s for {nsteps} steps)",
              file=sys.stderr, flush=True)
    (t1, v1, d1), (tn, vn, dn) = legs
    model = cpu_model()
    return {"value": vn, "unit": "cell-cycles/s", "cores": tn, "kind": "port",
            "sample": f"{sample}; C++ oracle with OpenMP over (j, k) rows, {tn} threads, {dn:.1f} s; host {model}, "
                      f"{os.cpu_count()} logical cpus",
            "single_thread": {"value": v1, "cores": 1, "seconds": d1}, "cpu_model": model}


def roofline(st, dom, nl, ms, nsteps, iters, workload, folded=False):
    avg_s = ms / nl / 1e3
    byts = kernel_bytes(st, dom, nl, nsteps, {"iters_per_solve": iters / nsteps, "folded": folded}) / nl
    flops = kernel_flops(st, dom, nl, folded) / nl
    peak = FP32_MFMA_PEAK if dom in ("sep_transform_f32", "sep_modes") else FP64_PEAK
    if byts > 0 and (flops == 0 or flops / byts < peak / HBM_PEAK):
        roof = {"bound": "hbm", "achieved": byts / avg_s / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                "frac": byts / avg_s / HBM_PEAK, "traffic": None, "kernel": dom, "bytes_per_launch": byts,
                "avg_launch_us": avg_s * 1e6,
                "note": "algorithmic bytes: each array the launch must touch counted once (FP64 node-shaped vectors, "
                        "FP32 quarters); durations from HIP events recorded at each launch's start and end over the "
                        "same number of cycles as the timed region"}
    else:
        roof = {"bound": "mfma", "achieved": flops / avg_s / 1e12, "peak": peak / 1e12, "unit": "TFLOP/s",
                "frac": flops / avg_s / peak, "traffic": None, "kernel": dom, "flops_per_launch": flops,
                "bytes_per_launch": byts, "avg_launch_us": avg_s * 1e6,
                "note": (("x-y transform launch of the Float32 preconditioner's twelve parity quarters on "
                          "v_mfma_f32_16x16x4f32 (half-size bases, 2 (ceil(Lx/2) + ceil(Ly/2)) flop per quarter "
                          "element)" if folded else
                          "fused x-y transform launch of the Float32 preconditioner on v_mfma_f32_16x16x4f32 "
                          "(2 (Lx + Ly) flop per element)")
                         if dom == "sep_transform_f32" else
                         "fused x-y transform launch of the FP64 separable solve (2 (Lx + Ly) flop per element)")
                        + "; durations from HIP events recorded at each launch's start and end over the same "
                          "number of cycles as the timed region"}
    roof["traffic"] = measured_traffic(dom, workload)
    return roof


def run_workload(name, steps, warmup, profile=True, cpu_steps=0, profile_out=None, tune=False):
    """Warm-up and timed ladder cycles of one workload with its state resident on the device; then,
    with `profile`, the same number of cycles again with HIP events around every launch (the
    roofline kernel), and with `cpu_steps` the CPU legs.  With `tune`, the library's tuner picks the
    launch shapes on the seeded state first (untimed; the state is left as it was)."""
    make, desc, mask = WORKLOADS[name]
    st, p = make()
    H, HJ = p["H"], st.selffield(p["J"])
    s0 = seeded_state(st.model, st.grid, st.masks, H, seed=20261202)
    dev = HipStepper(st)
    dev.upload(s0)
    print(f"[bench] {name}: dt={st.dt:.5f} stages={st.rkl} nsub={st.nsub} grid={st.grid.Nx}x{st.grid.Ny}x"
          f"{st.grid.Nz} padded", file=sys.stderr, flush=True)
    tuned = None
    if tune and st.field_mode == 2:
        tuned = dev.autotune(H, HJ)
        print(f"[bench] {name}: tuned box_kc={tuned['box_kc']} plane_mt={tuned['plane_mt']} "
              f"zmodes_waves={tuned['zmodes_waves']} qparts={tuned['qparts']} fold={tuned['fold']} zcol={tuned['zcol']}: " +
              ", ".join(f"{k}={v}:{us:.1f}us" for k, v, us in tuned["candidates"]),
              file=sys.stderr, flush=True)
    for q in range(warmup):
        dev.step(LMAX, H, HJ)
        print(f"[bench] warmup cycle {q + 1}/{warmup} done", file=sys.stderr, flush=True)
    st0 = dev.stats()
    t = time.perf_counter()
    for q in range(steps):
        dev.step(LMAX, H, HJ)
    wall = time.perf_counter() - t
    st1 = dev.stats()
    chk = dev.download()
    gmax = max(float(np.abs(g).max()) for g in chk.G)
    if not np.isfinite(gmax) or gmax > 5:
        raise RuntimeError(f"{name}: the state diverged (max|psi| = {gmax})")
    dev_ms = st1["device_ms"] - st0["device_ms"]
    iters = st1["pcg_iters"] - st0["pcg_iters"]
    rollbacks = st1["rollbacks"] - st0["rollbacks"]
    captures = st1["captures"] - st0["captures"]
    nsteps = steps * LMAX
    print(f"[bench] {name}: timed {steps} cycles: wall {wall:.3f}s device {dev_ms / 1e3:.3f}s, pcg iters/step "
          f"{iters / nsteps:.2f}, roll backs {rollbacks}, graph captures {captures}", file=sys.stderr, flush=True)
    cells = st.grid.cells
    value = cells * steps / wall
    nodes = (st.grid.Nx + 1) * (st.grid.Ny + 1) * (st.grid.Nz + 1)
    balg = cycle_bytes_per_node_step(st, iters / nsteps)
    out = {"value": value, "wall": wall, "steps": steps, "st": st, "nsteps": nsteps, "iters": iters,
           "config": {"workload": desc, "cells": cells, "lmax": LMAX, "pcg_iters_per_step": iters / nsteps,
                      "device_ms_per_cycle": dev_ms / steps, "padded_nodes": nodes, "cell_steps_per_s": value * LMAX,
                      "b_alg_per_node_step": balg,
                      "hbm_roofline_frac_cycle": balg * nodes * LMAX * steps / wall / HBM_PEAK,
                      "ms_per_time_step": 1e3 * wall / nsteps, "timed_rollbacks": rollbacks,
                      "timed_graph_captures": captures, "pcg_budget": st1.get("pcg_budget"),
                      "tuned": None if tuned is None else {k: tuned[k] for k in ("box_kc", "plane_mt", "zmodes_waves",
                                                                                 "qparts", "fold", "zcol")},
                      **({"pcg_refine": {"inner_tol": st.pcg_inner_tol, "rounds": st.pcg_rounds}}
                         if getattr(st, "pcg_refine", False) else {})},
           "mask": mask, "roofline": None, "cpu": None}
    if profile:
        dev.profile(True)
        for q in range(steps):
            dev.step(LMAX, H, HJ)
        prof = dev.profile_get()
        dev.profile(False)
        tot = sum(ms for _, ms in prof.values())
        rows = sorted(prof.items(), key=lambda kv: -kv[1][1])
        for kname, (nl, ms) in rows:
            if nl:
                print(f"[bench]   {kname:20s} {nl:8d} launches {ms:10.3f} ms ({100 * ms / tot:5.1f}%) "
                      f"{1e3 * ms / nl:8.2f} us/launch", file=sys.stderr)
        dom, (nl, ms) = rows[0]
        out["roofline"] = roofline(st, dom, nl, ms, nsteps, iters, name, bool(st1.get("folded", 0)))
        out["config"]["folded_preconditioner"] = bool(st1.get("folded", 0))
        out["config"]["kernel_ms_per_cycle"] = {k: v[1] / steps for k, v in prof.items() if v[0]}
        # every class with algorithmic bytes: its average launch against the HBM peak (and the
        # transforms against the matrix peak)
        per = {}
        for kname, (nl_k, ms_k) in prof.items():
            if not nl_k:
                continue
            b = kernel_bytes(st, kname, nl_k, nsteps, {"iters_per_solve": iters / nsteps,
                                                       "folded": bool(st1.get("folded", 0))}) / nl_k
            if b <= 0:
                continue
            us = 1e3 * ms_k / nl_k
            row = {"us_per_launch": us, "bytes_per_launch": b, "GB_s": b / us / 1e3, "hbm_frac": b / us / 1e3 / (HBM_PEAK / 1e9)}
            f = kernel_flops(st, kname, nl_k, bool(st1.get("folded", 0))) / nl_k
            if f > 0:
                pk = FP32_MFMA_PEAK if kname in ("sep_transform_f32", "sep_modes") else FP64_PEAK
                row["TFLOP_s"] = f / us / 1e6
                row["mfma_frac"] = f / us / 1e6 / (pk / 1e12)
            per[kname] = row
        out["config"]["kernel_roofline"] = per
        if profile_out:
            with open(profile_out, "w") as fh:
                json.dump({k: {"launches": v[0], "ms": v[1]} for k, v in prof.items()}, fh, indent=1)
    if cpu_steps:
        out["cpu"] = cpu_sample(name, st, dev, H, HJ, cpu_steps)
    dev.close()
    return out


def bench_serve(args) -> int:
    """The served current sweep: requests (juliatdgl_amd.sweep.serving_specs: two or three bands, grids
    log-uniform from 24x16x4 to 256x128x32, a notch or a hole, a random current and polarity, thermal
    noise T~ = 1e-3 drawn on the device) admitted by HBM footprint as slots free up, each run ending
    at its first phase slip across the voltmeter (crossing_times with stop = 1) or at t = T.  First the
    stream's capacity (every request there at t = 0: runs/hour), then a Poisson stream at
    `serve_load` of that rate (or at --serve-rate requests/s): its p50/p99 latency (finish - arrival)
    and whether every result equals the run's own alone (the first phase slip's time)."""
    from juliatdgl_amd import sweep
    n = args.serve_runs
    specs = sweep.serving_specs(n, master=4242)
    cells = sum(s.nx * s.ny * s.nz for s in specs)
    work = sweep.FirstSlip(T=args.serve_T, every=10)
    print(f"[bench] serve: {n} requests, {cells} cells, first slip or t = {args.serve_T}", file=sys.stderr, flush=True)
    sweep.run_batch(specs[:2], sweep.Cycles(cycles=1), workers=2)  # the device and the library warm
    srv = sweep.Server(budget_bytes=200e9, slots=args.workers)
    for s in specs:
        srv.submit(s, 0.0)
    t0 = time.perf_counter()
    full = srv.run(work)
    t_cap = time.perf_counter() - t0
    cap = n / t_cap
    print(f"[bench] capacity: {n} runs in {t_cap:.2f} s ({3600 * cap:.0f} runs/hr)", file=sys.stderr, flush=True)
    rate = args.serve_rate if args.serve_rate > 0 else args.serve_load * cap
    srv = sweep.Server(budget_bytes=200e9, slots=args.workers)
    arrivals = sweep.poisson_arrivals(n, rate=rate, seed=11)
    for s, t in zip(specs, arrivals):
        srv.submit(s, t)
    t0 = time.perf_counter()
    done = srv.run(work)
    wall = time.perf_counter() - t0
    lat = sweep.percentiles(d["latency"] for d in done.values())
    slipped = sum(1 for d in done.values() if d["result"][0] is not None)
    same = all(done[q]["result"] == full[q]["result"] for q in done)
    print(f"[bench] Poisson at {rate:.2f}/s: {n} runs in {wall:.2f} s: p50 {lat['p50']:.3f} s, p99 {lat['p99']:.3f} s, "
          f"{slipped} slipped, results as at capacity: {same}", file=sys.stderr, flush=True)
    line = {"metric": "served runs/hr", "value": 3600.0 * cap, "unit": "runs/hr", "n_gpus": 1, "steps": n,
            "warmup": 1, "ms_per_step": 1e3 * t_cap / n, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded psi per run, thermal noise from the device RNG)",
            "config": {"workload": "served current sweep: two- and three-band runs, grids log-uniform 24x16x4..256x128x32, "
                                   "notch or hole, |J| in [0.3, 1.5] of either polarity, noise T~ = 1e-3, separable "
                                   f"field, dt 0.01, each to its first phase slip or t = {args.serve_T}; value: every "
                                   "request at t = 0 (capacity); latency: a Poisson stream below it",
                       "runs": n, "cells": cells, "slots": args.workers, "capacity_s": t_cap,
                       "poisson_rate_per_s": rate, "poisson_wall_s": wall, "poisson_runs_per_hr": 3600.0 * n / wall,
                       "p50_latency_s": lat["p50"], "p99_latency_s": lat["p99"], "mean_latency_s": lat["mean"],
                       "slipped": slipped, "same_as_at_capacity": same,
                       "bands": {str(b): sum(1 for s in specs if s.bands == b) for b in (2, 3)}},
            "roofline": None, "cpu_baseline": None}
    print(json.dumps(line), flush=True)
    return 0


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10, help="timed ladder cycles")
    ap.add_argument("--warmup", type=int, default=5,
                    help="untimed ladder cycles (the PCG budget settles and its graph is captured in them)")
    ap.add_argument("--workload", default="open_3band_256x128x32", choices=tuple(WORKLOADS) + ("sweep64", "serve"))
    ap.add_argument("--serve-runs", type=int, default=64, help="serve: requests of the stream")
    ap.add_argument("--serve-rate", type=float, default=0.0,
                    help="serve: Poisson arrival rate (requests/s); 0: --serve-load of the measured capacity")
    ap.add_argument("--serve-load", type=float, default=0.8, help="serve: the Poisson rate as a fraction of capacity")
    ap.add_argument("--serve-T", type=float, default=40.0, help="serve: a run ends at its first slip or at this time")
    ap.add_argument("--workers", type=int, default=8, help="sweep64: runs in flight at once")
    ap.add_argument("--cpu-steps", type=int, default=0,
                    help="time steps of the CPU baseline sample (default 8)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-secondary", action="store_true",
                    help="leave out the 128x64x16 KWT workload that the 256x128x32 line carries beside it")
    ap.add_argument("--inner-tol", type=float, default=1e-4,
                    help="open_3band_256x128x32_refine: the inner solves' relative tolerance")
    ap.add_argument("--rounds", type=int, default=3, help="open_3band_256x128x32_refine: rounds per captured solve")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the library's default launch shapes (by default its tuner picks them before the warm-up)")
    ap.add_argument("--profile-out", default=None, help="write the per-kernel profile (JSON) here")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="first copy a MIB-MiB device buffer (a reference kernel for the counters' scale)")
    args = ap.parse_args(argv)
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: a run is one device's work and this build has no multi-GPU path "
              "(--gpus 1 only)", file=sys.stderr)
        return 2
    if device_count() < 1:
        print("no HIP device", file=sys.stderr)
        return 1
    if args.workload == "sweep64":
        return bench_sweep(args)
    if args.workload == "serve":
        return bench_serve(args)
    REFINE.update(pcg_inner_tol=args.inner_tol, pcg_rounds=args.rounds)
    if args.copy_probe:
        copy_probe(args.copy_probe)
    cpu_steps = 0 if args.no_cpu else (args.cpu_steps or 8)
    r = run_workload(args.workload, args.steps, args.warmup, profile=True, cpu_steps=cpu_steps, tune=not args.no_tune,
                     profile_out=args.profile_out)
    published = PUBLISHED.get(args.workload)
    cfg = dict(r["config"])
    if published:
        cfg["reference_published"] = published
    if args.workload == "open_3band_256x128x32" and not args.no_secondary:
        sec = run_workload("kwt_pcg_128x64x16", min(args.steps, 10), 3, profile=False, tune=not args.no_tune)
        cfg["secondary"] = {"kwt_pcg_128x64x16": {"cell_cycles_per_s": sec["value"],
                                                  "ms_per_cycle": 1e3 * sec["wall"] / sec["steps"],
                                                  **{k: sec["config"][k] for k in ("workload", "cells",
                                                                                    "pcg_iters_per_step",
                                                                                    "hbm_roofline_frac_cycle")}}}
    value = r["value"]
    st = r["st"]
    out = {"metric": "cell-cycles/s", "value": value, "unit": "cell-cycles/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": 1e3 * r["wall"] / args.steps, "higher_is_better": True,
           "scaling": "weak",
           "vs_baseline": (value / (st.grid.cells / (published["t4_cycle_ms_per_step"] * 1e-3) / LMAX)) if published else None,
           "dtype": "f64", "data": "synthetic (seeded psi = 1 + 0.01 (n1 + i n2)), " + r["mask"],
           "config": cfg, "roofline": r["roofline"], "cpu_baseline": r["cpu"]}
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
