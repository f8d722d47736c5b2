"""Ladder-cycle throughput of the HIP stepper on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload kwt_pcg_128x64x16]

A "step" of this benchmark is one ladder cycle: lMax = 40 time steps of the release-2.0 step
(the n that TDGL's ladder hands to its stepper hook, gtdgl/src/protocol.jl:28, 409), with the
state resident on the device.  The workload is the two-band KWT stepper on 128x64x16 cells with
an insulating buffer and the PCG field solve to relative residual 1e-10 (juliatdgl_amd.configs.kwt_pcg_128x64x16).
Prints one JSON line: cell-cycles/s (nx·ny·nz × cycles / s), the roofline of the dominant
kernel (HIP events around its launches over the timed cycles), and the CPU oracle timed on a
bounded sample of the same workload on this host.
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
        # per PCG solve of `it` iterations: it + 1 forward launches (FP64 in, FP32 out: 12 B per
        # element), the first inverse launch (FP32 in, FP64 out, r read for r·z: 20 B) and it
        # inverse launches of the flexible iteration (r and z_old read: 28 B)
        it = stats.get("iters_per_solve", 0.0)
        per = (12.0 * (it + 1) + 20.0 + 28.0 * it) / (2.0 * (it + 1))
        return per * rng * launches
    if name == "pcg_apply_A":
        return 24.0 * rng * launches
    if name == "current_field":
        return (16 * N + 48 + 24 * 6 + 24 + 24) * nodes * launches
    if name == "links_curl_bc":
        return 96.0 * nodes * launches
    return 0.0


def ranges(gr):
    return [(gr.Nx - 2, gr.Ny - 1, gr.Nz - 1), (gr.Nx - 1, gr.Ny - 2, gr.Nz - 1), (gr.Nx - 1, gr.Ny - 1, gr.Nz - 2)]


def kernel_flops(st, name: str, launches: int) -> float:
    """Algorithmic flops of the launches of one kernel class: a fused x-y transform launch of the
    separable solve multiplies every element by an (Lx x Lx) then an (Ly x Ly) matrix, 2 (Lx + Ly)
    flop per element."""
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
    pcg = 300 + 288 * iters_per_step if st.field_mode == 2 else (168 if st.field_mode == 1 else 0)
    return band + field + pcg + 96


def cpu_baseline(st, s0, H, HJ, nsteps: int, threads: int):
    """The oracle (CPU restatement of the step) on `nsteps` time steps of the same workload."""
    from oracle.oracle import Oracle, set_threads
    set_threads(threads)
    o = Oracle(st)
    o.set_state(s0)
    t = time.perf_counter()
    o.step(nsteps, H, HJ)
    dt = time.perf_counter() - t
    cells = st.grid.cells
    return cells * (nsteps / LMAX) / dt, dt


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
    print(f"[bench] batched kernels (one launch per kernel over all runs): {t_kb:.2f} s", file=sys.stderr, flush=True)
    kb16, t_kb16 = sweep.run_batched(specs, work, group=16, workers=args.workers)
    print(f"[bench] batched kernels, groups of 16 set up while the last runs: {t_kb16:.2f} s", file=sys.stderr, flush=True)
    same = all(seq[q] == bat[q] == kb[q] == kb16[q] for q in seq)
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
                       "cell_cycles_per_s_batched": cells * args.steps / t_bat,
                       "same_alone_and_in_the_mix": bool(same and same_srv),
                       "served_p50_latency_s": lat["p50"], "served_p99_latency_s": lat["p99"],
                       "served_arrival_rate_per_s": 0.9 * rate, "served_runs_per_hr": 3600.0 * len(specs) / t_srv},
            "roofline": None, "cpu_baseline": None}
    print(json.dumps(line), flush=True)
    return 0


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10, help="timed ladder cycles")
    ap.add_argument("--warmup", type=int, default=5,
                    help="untimed ladder cycles (the PCG budget settles and its graph is captured in them)")
    ap.add_argument("--workload", default="kwt_pcg_128x64x16",
                    choices=("kwt_pcg_128x64x16", "open_3band_256x128x32", "pcg_wedge_96x48x24", "plain_wedge_96x48x24",
                             "sweep64"))
    ap.add_argument("--workers", type=int, default=8, help="sweep64: runs in flight at once")
    ap.add_argument("--cpu-steps", type=int, default=8, help="time steps of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
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
    if args.workload == "kwt_pcg_128x64x16":
        st, p = configs.kwt_pcg_128x64x16()
        workload = ("two-band KWT (u=Gamma/sigma_n=5.79, gkwt=10) TDGL, 128x64x16 cells, dx=0.5, buf 2 insulating "
                    "(sigma_vac/sigma_n=0.01), PCG tol 1e-10 (Float32 preconditioner, FP64 outer), 30-degree wedge, "
                    "H=0.2, J=0.1, dt_scale=0.1, ladder cycle = 40 steps")
    elif args.workload == "open_3band_256x128x32":
        st, p = configs.open_3band_256x128x32()
        workload = ("three-band anisotropic (m_z = 4 m_xy) TDGL, 256x128x32 cells, dx=0.5, buf 8 insulating "
                    "(sigma_vac/sigma_n=0.01), PCG tol 1e-8 (Float32 preconditioner, FP64 outer), open boundary "
                    "(FFT convolution, every t_GL), notch, H=0.2, J=0.1, ladder cycle = 40 steps")
    elif args.workload == "pcg_wedge_96x48x24":
        st, p = configs.pcg_wedge_96x48x24()
        workload = ("the reference's validation run file pcg96.toml: mgb2_model_II at kappa=2, 96x48x24 cells, dx=0.5, "
                    "buf 2 insulating (sigma_vac=0.01), PCG tol 1e-8 (Float32 preconditioner, FP64 outer), WEDGE, "
                    "J=0.3, H=0.03, ladder cycle = 40 steps")
    elif args.workload == "plain_wedge_96x48x24":
        st, p = configs.plain_wedge_96x48x24()
        workload = ("the reference's validation run file plain96.toml: mgb2_model_II at kappa=2, 96x48x24 cells, "
                    "dx=0.5, buf 1, separable implicit field solve, WEDGE, J=0.3, H=0.03, ladder cycle = 40 steps")
    else:
        raise SystemExit(f"unknown workload {args.workload}")
    # the reference's published T4 figures for these run files (ms per time step, median of three
    # batches of 20 steps; gpu/TDGLGPU/validation/README.md:41-71)
    published = {"pcg_wedge_96x48x24": {"t4_batch_ms_per_step": 26.903, "t4_cycle_ms_per_step": 21.040,
                                        "laptop_1core_ms_per_step": 146.370},
                 "plain_wedge_96x48x24": {"t4_batch_ms_per_step": 7.295, "t4_cycle_ms_per_step": 4.816,
                                          "laptop_1core_ms_per_step": 22.574}}.get(args.workload)
    if args.copy_probe:
        copy_probe(args.copy_probe)
    H, HJ = p["H"], st.selffield(p["J"])
    s0 = seeded_state(st.model, st.grid, st.masks, H, seed=20261202)
    dev = HipStepper(st)
    dev.upload(s0)
    print(f"[bench] {args.workload}: dt={st.dt:.5f} stages={st.rkl} nsub={st.nsub} grid={st.grid.Nx}x{st.grid.Ny}x"
          f"{st.grid.Nz} padded", file=sys.stderr, flush=True)
    for q in range(args.warmup):
        dev.step(LMAX, H, HJ)
        print(f"[bench] warmup cycle {q + 1}/{args.warmup} done", file=sys.stderr, flush=True)
    st0 = dev.stats()
    t = time.perf_counter()
    for q in range(args.steps):
        dev.step(LMAX, H, HJ)
    wall = time.perf_counter() - t
    st1 = dev.stats()
    chk = dev.download()
    gmax = max(float(np.abs(g).max()) for g in chk.G)
    if not np.isfinite(gmax) or gmax > 5:
        print(f"[bench] the state diverged (max|psi| = {gmax})", file=sys.stderr)
        return 3
    dev_ms = st1["device_ms"] - st0["device_ms"]
    iters = st1["pcg_iters"] - st0["pcg_iters"]
    rollbacks = st1["rollbacks"] - st0["rollbacks"]
    captures = st1["captures"] - st0["captures"]
    nsteps = args.steps * LMAX
    print(f"[bench] timed {args.steps} cycles: wall {wall:.3f}s device {dev_ms / 1e3:.3f}s, pcg iters/step "
          f"{iters / nsteps:.2f}, roll backs {rollbacks}, graph captures {captures}", file=sys.stderr, flush=True)
    cells = st.grid.cells
    value = cells * args.steps / wall
    # per-kernel device time: the same number of cycles again with HIP events around every launch
    dev.profile(True)
    for q in range(args.steps):
        dev.step(LMAX, H, HJ)
    prof = dev.profile_get()
    dev.profile(False)
    tot = sum(ms for _, ms in prof.values())
    rows = sorted(prof.items(), key=lambda kv: -kv[1][1])
    for name, (nl, ms) in rows:
        if nl:
            print(f"[bench]   {name:20s} {nl:8d} launches {ms:10.3f} ms ({100 * ms / tot:5.1f}%) "
                  f"{1e3 * ms / nl:8.2f} us/launch", file=sys.stderr)
    dom, (nl, ms) = rows[0]
    avg_s = ms / nl / 1e3
    byts = kernel_bytes(st, dom, nl, nsteps, {"iters_per_solve": iters / nsteps}) / nl
    flops = kernel_flops(st, dom, nl) / nl
    peak = FP32_MFMA_PEAK if dom == "sep_transform_f32" else FP64_PEAK
    if byts > 0 and (flops == 0 or flops / byts < peak / HBM_PEAK):
        roof = {"bound": "hbm", "achieved": byts / avg_s / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                "frac": byts / avg_s / HBM_PEAK, "traffic": None, "kernel": dom, "bytes_per_launch": byts,
                "avg_launch_us": avg_s * 1e6}
    else:
        roof = {"bound": "mfma", "achieved": flops / avg_s / 1e12, "peak": peak / 1e12, "unit": "TFLOP/s",
                "frac": flops / avg_s / peak, "traffic": None, "kernel": dom, "flops_per_launch": flops,
                "bytes_per_launch": byts, "avg_launch_us": avg_s * 1e6,
                "note": ("fused x-y transform launch of the Float32 preconditioner on v_mfma_f32_16x16x4f32"
                         if dom == "sep_transform_f32" else
                         "fused x-y transform launch of the FP64 separable solve on the vector ALUs")
                        + " (2 (Lx + Ly) flop per element); durations from HIP events recorded at each launch's "
                          "start and end over the same number of cycles as the timed region"}
    traffic = measured_traffic(dom, args.workload)
    roof["traffic"] = traffic
    if args.profile_out:
        with open(args.profile_out, "w") as fh:
            json.dump({k: {"launches": v[0], "ms": v[1]} for k, v in prof.items()}, fh, indent=1)
    cpu = None
    if not args.no_cpu:
        s1 = dev.download()
        v1, t1 = cpu_baseline(st, s1, H, HJ, args.cpu_steps, 1)
        cpu = {"value": v1, "unit": "cell-cycles/s", "cores": 1, "kind": "port",
               "sample": f"{args.cpu_steps} time steps (of 40 per cycle) of the same workload from the GPU's state "
                         f"after warmup, C++ oracle single-threaded, {t1:.1f} s; host {os.cpu_count()} logical cpus"}
        print(f"[bench] cpu oracle 1 thread: {v1:.3e} cell-cycles/s ({t1:.1f}s for {args.cpu_steps} steps)",
              file=sys.stderr, flush=True)
    nodes = (st.grid.Nx + 1) * (st.grid.Ny + 1) * (st.grid.Nz + 1)
    balg = cycle_bytes_per_node_step(st, iters / nsteps)
    hbm_cycle = balg * nodes * LMAX * args.steps / wall / HBM_PEAK
    out = {"metric": "cell-cycles/s", "value": value, "unit": "cell-cycles/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": 1e3 * wall / args.steps, "higher_is_better": True,
           "scaling": "weak",
           "vs_baseline": (value / (st.grid.cells / (published["t4_cycle_ms_per_step"] * 1e-3) / LMAX)) if published else None, "dtype": "f64",
           "data": "synthetic (seeded psi = 1 + 0.01 (n1 + i n2)), " + {"kwt_pcg_128x64x16": "30-degree sloped wedge mask",
                                                                         "open_3band_256x128x32": "notch mask"}.get(
               args.workload, "catalogue WEDGE mask"),
           "config": {"workload": workload, "cells": cells, "lmax": LMAX, "pcg_iters_per_step": iters / nsteps,
                      "device_ms_per_cycle": dev_ms / args.steps, "padded_nodes": nodes,
                      "cell_steps_per_s": value * LMAX, "b_alg_per_node_step": balg,
                      "hbm_roofline_frac_cycle": hbm_cycle, "ms_per_time_step": 1e3 * wall / nsteps,
                      "timed_rollbacks": rollbacks, "timed_graph_captures": captures,
                      **({"reference_published": published} if published else {})},
           "roofline": roof, "cpu_baseline": cpu}
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
