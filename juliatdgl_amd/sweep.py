"""Many runs on one device: a current sweep advanced together, and a request loop that admits runs
by their HBM footprint as others finish.

A ladder's runs are independent (the two polarities of run_ladders, protocol.jl:842-861, the points
of a critical-current sweep, the requests of a server), and one run of a few hundred thousand
nodes leaves most of an MI355X idle: its kernels are latency-bound.  Here every run keeps its own
device stepper (state, PCG history, noise stream resident in HBM) on its own HIP stream, and host
threads drive several runs at once, so the kernels of different runs overlap on the device's
hardware queues.  A run's result depends on its own description only (its seed, its stream of
normals, its deterministic reductions): the same alone and in the mix.

    specs = sweep_specs()                                   # 32 currents x 2 polarities, seeded grids and notches
    res = run_batch(specs, work=Cycles(40), workers=8)      # all of them, 8 at a time
    srv = Server(budget_bytes=200e9, slots=8); srv.submit(spec, t_arrival) ...; srv.run(work)
"""
from __future__ import annotations

import contextlib
import heapq
import math
import threading
import time
from dataclasses import dataclass, field
from typing import Callable, List, Optional

import numpy as np

from .configs import three_band_aniso_model, two_band_model
from .scheme import Grid, Stepper, catalogue_masks, hole_masks, make_stepper
from .state import State, seeded_state

MASK64 = (1 << 64) - 1


def splitmix64(x: int):
    """One SplitMix64 step (Steele, Lea & Flood, OOPSLA 2014): (next state, output)."""
    x = (x + 0x9E3779B97F4A7C15) & MASK64
    z = x
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
    return x, z ^ (z >> 31)


def run_seed(master: int, run_id: int) -> int:
    """The seed of run `run_id` under the sweep's master seed."""
    _, v = splitmix64((master ^ ((run_id * 0xD1B54A32D192ED03) & MASK64)) & MASK64)
    return v


@dataclass(frozen=True)
class RunSpec:
    """One run: the sample (nx, ny, nz cells, dx 0.5, a notch of depth `depth`·ny on face y1 or a
    hole of radius `hole` cells), the current J (its sign the polarity), the field H, its seed,
    the number of bands (2: the two bands of the 12x8x4 notch case; 3: three anisotropic bands,
    a = 1, 0.6, 0.3, ξ = 1, 1.8, 3, m_z = 4 m_xy), and optionally thermal noise T̃, its normals
    drawn on the device (Philox keyed by the seed) or, with noise_host, on the host from the
    run's own seeded stream and handed over before each call (the validation mode)."""
    run_id: int
    nx: int
    ny: int
    nz: int
    J: float
    depth: float = 0.35
    hole: float = 0.0
    H: float = 0.3
    seed: int = 1
    noise: Optional[float] = None
    bands: int = 2
    noise_host: bool = False


def sweep_specs(n_currents: int = 32, master: int = 2026) -> List[RunSpec]:
    """The batched sweep: `n_currents` currents in [0.05, 0.4] times both polarities; per run a grid
    drawn uniformly on multiples of 4 in [32,160]x[16,96]x[4,24] and a notch depth in [0.1, 0.4] of
    the width, from the run's own seed."""
    Js = np.linspace(0.05, 0.4, n_currents)
    out = []
    for q in range(2 * n_currents):
        s = run_seed(master, q)
        rng = np.random.default_rng(s)
        nx = 4 * int(rng.integers(8, 41))
        ny = 4 * int(rng.integers(4, 25))
        nz = 4 * int(rng.integers(1, 7))
        depth = float(rng.uniform(0.1, 0.4))
        J = float(Js[q // 2]) * (1.0 if q % 2 == 0 else -1.0)
        out.append(RunSpec(q, nx, ny, nz, J, depth=depth, seed=s & 0x7FFFFFFF))
    return out


def serving_specs(n: int, master: int = 4242, lo=(24, 16, 4), hi=(256, 128, 32), noise: float = 1e-3,
                  noise_host: bool = False) -> List[RunSpec]:
    """The requests of the served current sweep: per request, from its own seed, two or three bands,
    cells per axis drawn log-uniform between `lo` and `hi` (rounded to a multiple of 4 in x and y,
    of 2 in z), a notch (depth 0.1..0.4 of the width) or a hole (radius 0.1..0.25 of the smaller
    in-plane side), a current |J| in [0.3, 1.5] with a random polarity, and thermal noise T̃ (the
    noise temperature of the functional; the bath T sits in the bands' a)."""
    out = []
    llo, lhi = np.log(np.asarray(lo, float)), np.log(np.asarray(hi, float))
    for q in range(n):
        s = run_seed(master, q)
        rng = np.random.default_rng(s)
        d = np.exp(rng.uniform(llo, lhi))
        nx = int(min(hi[0], max(lo[0], 4 * round(d[0] / 4))))
        ny = int(min(hi[1], max(lo[1], 4 * round(d[1] / 4))))
        nz = int(min(hi[2], max(lo[2], 2 * round(d[2] / 2))))
        bands = 2 + int(rng.integers(0, 2))
        hole = float(rng.uniform(0.1, 0.25) * min(nx, ny)) if rng.random() < 0.5 else 0.0
        depth = float(rng.uniform(0.1, 0.4))
        J = float(rng.uniform(0.3, 1.5)) * (1.0 if rng.random() < 0.5 else -1.0)
        out.append(RunSpec(q, nx, ny, nz, J, depth=depth, hole=hole, seed=s & 0x7FFFFFFF, noise=noise, bands=bands,
                           noise_host=noise_host))
    return out


def model_of(spec: RunSpec):
    if spec.bands == 2:
        return two_band_model()
    if spec.bands == 3:
        return three_band_aniso_model()
    raise ValueError(f"a run has 2 or 3 bands, not {spec.bands}")


def stepper_of(spec: RunSpec) -> Stepper:
    """The plain FP64 stepper of a run: its bands (κ = 2.5), the separable implicit field, the bar
    self-field, dt = 0.01 (Euler substeps or stages as the bands need)."""
    md = model_of(spec)
    gr = Grid.make(spec.nx, spec.ny, spec.nz, dx=0.5, buf=1)
    if spec.hole > 0:
        mk = hole_masks(gr, (spec.nx / 2, spec.ny / 2), spec.hole)
    else:
        mk = catalogue_masks(gr, "WEDGE", wedge_depth=spec.depth)
    kw = {}
    if spec.noise is not None:
        kw = dict(noise=spec.noise, noise_seed=spec.seed, noise_draws="host" if spec.noise_host else "device")
    return make_stepper(md, gr, mk, dt=0.01, selffield_bc="bar", **kw)


class HostDraws:
    """A backend (HipStepper, or the oracle's) whose steps take their normals from the run's own
    seeded stream on the host (numpy PCG64 of the run's seed), handed over before every call:
    the validation mode of a noisy run, in which the device and the oracle see the same draws."""

    def __init__(self, backend, st: Stepper, seed: int):
        from .noise import per_step
        self.b, self.st = backend, st
        self.rng = np.random.default_rng(seed)
        self.n = per_step(st)

    def step(self, n: int, H: float, HJ: float) -> None:
        from .noise import from_stream, pack
        if n > 0:
            self.b.noise_draws(pack(self.st, [from_stream(self.st, self.rng.standard_normal(self.n))
                                              for _ in range(n)]))
        self.b.step(n, H, HJ)

    def __getattr__(self, name):
        return getattr(self.b, name)


def state_of(spec: RunSpec, st: Stepper) -> State:
    return seeded_state(st.model, st.grid, st.masks, spec.H, seed=spec.seed)


def footprint(st: Stepper) -> int:
    """An upper bound of the device bytes of a run (tdgl_create's allocations): the node-shaped
    arrays of the state, the work arrays and the solve, at the padded pitch."""
    gr = st.grid
    px = (gr.Nx + 1 + 7) // 8 * 8
    n = px * (gr.Ny + 1) * (gr.Nz + 1)
    N = st.N
    arrays = (2 * 2 * N      # G, F
              + 3 * 2        # θ, φ
              + 3 * 2 * 2    # U, V (complex)
              + 3 * 3        # J, h, τ
              + 1 + 3        # f, D
              + 2 * 2 + 1    # W, W2, Ω
              + 6 * 2        # separable intermediates (double and float) and sources
              + 12)          # PCG vectors
    if st.field_mode == 2:
        arrays += 30 + 2 * N  # the roll-back copies
    return int(8 * n * arrays * 1.25) + (8 << 20)


# ------------------------------------------------------------------------------ the work of a run
@dataclass
class Cycles:
    """A run that advances `cycles` ladder cycles of `lmax` steps at its current, measuring every
    cycle on the device (tdgl_meter_cycle): the result is the last record."""
    cycles: int = 40
    lmax: int = 40
    delta_v: int = 3

    def __call__(self, dev, st: Stepper, spec: RunSpec) -> dict:
        from .meter import Voltmeter
        HJ = st.selffield(spec.J)
        dev.meter_begin(Voltmeter.of_ladder(st, self.delta_v))
        rec = None
        for m in range(1, self.cycles + 1):
            dev.step(self.lmax, spec.H, HJ)
            rec = dev.meter_cycle(m == 1)
        return rec


@dataclass
class FirstSlip:
    """A run that ends at its first phase slip across the voltmeter (crossing_times with stop = 1,
    protocol.jl:259-296) or at time `T`: the result is (time of the slip or None, time reached)."""
    T: float = 200.0
    every: int = 10
    delta_v: int = 3

    def __call__(self, dev, st: Stepper, spec: RunSpec):
        from .meter import crossing_times
        times, t, div = crossing_times(dev, st, spec.J, spec.H, self.T, stop=1, delta_v=self.delta_v, every=self.every)
        return (times[0] if times else None, t, div)


def _one_blas_thread():
    """BLAS on one thread while host threads drive several runs: each run's set-up (the bar
    field's quadrature is two small matrix products) would otherwise start a thread team per
    call, and eight concurrent teams oversubscribe the host's cores many times over."""
    try:
        from threadpoolctl import threadpool_limits
    except ImportError:  # no control over the BLAS pool: leave it as it is
        return contextlib.nullcontext()
    return threadpool_limits(limits=1, user_api="blas")


def _execute(spec: RunSpec, work: Callable, st: Optional[Stepper] = None):
    from .hip import HipStepper
    st = st if st is not None else stepper_of(spec)
    dev = HipStepper(st)
    try:
        dev.upload(state_of(spec, st))
        out = work(HostDraws(dev, st, spec.seed) if spec.noise_host else dev, st, spec)
        return out, dev.stats()["device_bytes"]
    finally:
        dev.close()


def run_batch(specs: List[RunSpec], work: Callable, workers: int = 8):
    """Every run of `specs`, `workers` at a time (1: one after the other).  Returns
    {run_id: result} and the wall time."""
    results = {}
    lock = threading.Lock()
    queue = list(specs)
    errors = []

    def worker():
        while True:
            with lock:
                if not queue or errors:
                    return
                spec = queue.pop(0)
            try:
                out, _ = _execute(spec, work)
            except BaseException as e:  # noqa: BLE001 - handed to the caller
                with lock:
                    errors.append(e)
                return
            with lock:
                results[spec.run_id] = out

    with _one_blas_thread():
        t0 = time.perf_counter()
        ths = [threading.Thread(target=worker) for _ in range(max(1, workers))]
        for t in ths:
            t.start()
        for t in ths:
            t.join()
        wall = time.perf_counter() - t0
    if errors:
        raise errors[0]
    return results, wall


def run_batched(specs: List[RunSpec], work: "Cycles", group: int = 64, workers: int = 8, lanes: int = 1):
    """Every run of `specs` with batched kernels (tdgl_batch_*): groups of `group` runs advanced
    together, one launch per kernel of a step over the group's runs of each path of the separable
    solve (with or without the parity folds: one batch per path).  With lanes = 1 host threads
    set up the next group (stepper, state, upload, meter) while the current one runs; with
    lanes > 1 that many groups step at once, each batch on its own stream (their launches overlap
    on the device, so one batch's slowest blocks do not idle it), each lane setting up its next
    group as it finishes one.  Returns {run_id: the last cycle's record} and the wall time."""
    from concurrent.futures import ThreadPoolExecutor
    from .hip import HipBatch, HipStepper
    from .meter import Voltmeter

    if not isinstance(work, Cycles):
        raise ValueError("run_batched: the batched runs take a fixed number of cycles (Cycles)")

    def prep(spec):
        st = stepper_of(spec)
        dev = HipStepper(st)
        dev.upload(state_of(spec, st))
        dev.meter_begin(Voltmeter.of_ladder(st, work.delta_v))
        return spec, st, dev

    def advance_part(part):
        batch = HipBatch([dev for _, _, dev in part])
        try:
            H = np.array([sp.H for sp, _, _ in part])
            HJ = np.array([st.selffield(sp.J) for sp, st, _ in part])
            recs = None
            for m in range(1, work.cycles + 1):
                batch.step(work.lmax, H, HJ)
                recs = batch.meter_cycle(m == 1)
            return {sp.run_id: rec for (sp, _, _), rec in zip(part, recs or [None] * len(part))}
        finally:
            batch.close()

    def advance(ready):
        # one batch per path of the separable solve (the parity folds, which a handle takes on planes
        # too large for the full-length kernel, or not): a batch of one path launches what each run's
        # own step launches, so every record equals the run's alone
        out = {}
        try:
            parts = {}
            for r in ready:
                parts.setdefault(r[2].stats()["folded"], []).append(r)
            for part in parts.values():
                out.update(advance_part(part))
            return out
        finally:
            for _, _, dev in ready:
                dev.close()

    groups = [specs[q:q + max(1, group)] for q in range(0, len(specs), max(1, group))]
    results = {}
    with _one_blas_thread(), ThreadPoolExecutor(max_workers=max(1, workers)) as ex:
        t0 = time.perf_counter()
        if lanes <= 1:
            pending = [ex.submit(prep, sp) for sp in groups[0]] if groups else []
            for k in range(len(groups)):
                ready = [f.result() for f in pending]
                pending = [ex.submit(prep, sp) for sp in groups[k + 1]] if k + 1 < len(groups) else []
                results.update(advance(ready))
        else:
            lock = threading.Lock()
            queue = list(range(len(groups)))
            errors = []

            def lane():
                while True:
                    with lock:
                        if not queue or errors:
                            return
                        k = queue.pop(0)
                    try:
                        out = advance(list(ex.map(prep, groups[k])))
                    except BaseException as e:  # noqa: BLE001 - handed to the caller
                        with lock:
                            errors.append(e)
                        return
                    with lock:
                        results.update(out)

            ths = [threading.Thread(target=lane) for _ in range(lanes)]
            for t in ths:
                t.start()
            for t in ths:
                t.join()
            if errors:
                raise errors[0]
        wall = time.perf_counter() - t0
    return results, wall


@dataclass(order=True)
class _Req:
    t_arrival: float
    run_id: int
    spec: RunSpec = field(compare=False)


class Server:
    """A request loop with continuous admission: runs arrive at their times; a run is admitted as
    soon as a slot is free and its footprint fits in what the running ones leave of the HBM budget
    (first come, first served; a run that cannot fit waits for the ones ahead of it to finish);
    several admitted runs overlap on the device.  Latency = finish − arrival."""

    def __init__(self, budget_bytes: float = 200e9, slots: int = 8):
        self.budget = float(budget_bytes)
        self.slots = int(slots)
        self.pending: List[_Req] = []
        self.done = {}

    def submit(self, spec: RunSpec, t_arrival: float = 0.0) -> None:
        heapq.heappush(self.pending, _Req(float(t_arrival), spec.run_id, spec))

    def run(self, work: Callable):
        """Serve every submitted request; returns {run_id: dict(result, arrival, start, finish,
        latency, bytes)} with times in seconds from the start of the loop."""
        cv = threading.Condition()
        state = {"used": 0.0, "running": 0, "error": None}
        t0 = time.perf_counter()
        ths = []

        def launch(req, need, st):
            def body():
                start = time.perf_counter() - t0
                try:
                    out, nbytes = _execute(req.spec, work, st)
                except BaseException as e:  # noqa: BLE001
                    with cv:
                        state["error"] = e
                        state["used"] -= need
                        state["running"] -= 1
                        cv.notify_all()
                    return
                finish = time.perf_counter() - t0
                with cv:
                    self.done[req.run_id] = dict(result=out, arrival=req.t_arrival, start=start, finish=finish,
                                                 latency=finish - req.t_arrival, bytes=nbytes)
                    state["used"] -= need
                    state["running"] -= 1
                    cv.notify_all()
            th = threading.Thread(target=body)
            ths.append(th)
            th.start()

        steppers = {}
        with _one_blas_thread():
            while True:
                build = None
                with cv:
                    if not self.pending or state["error"] is not None:
                        break
                    req = self.pending[0]
                    now = time.perf_counter() - t0
                    if req.t_arrival > now:
                        cv.wait(timeout=req.t_arrival - now)
                        continue
                    if req.run_id not in steppers:
                        build = req  # set up outside the lock: finishing runs take it meanwhile
                    else:
                        st = steppers[req.run_id]
                        need = float(footprint(st))
                        if need > self.budget:
                            raise ValueError(f"run {req.run_id} needs {need:.3g} bytes, more than the budget")
                        if state["running"] >= self.slots or state["used"] + need > self.budget:
                            cv.wait(timeout=0.05)
                            continue
                        heapq.heappop(self.pending)
                        state["used"] += need
                        state["running"] += 1
                        launch(req, need, steppers.pop(req.run_id))
                if build is not None:
                    steppers[build.run_id] = stepper_of(build.spec)
            for th in ths:
                th.join()
        if state["error"] is not None:
            raise state["error"]
        return self.done


def poisson_arrivals(n: int, rate: float, seed: int = 7) -> List[float]:
    """n arrival times of a Poisson stream of `rate` requests per second."""
    rng = np.random.default_rng(seed)
    return list(np.cumsum(rng.exponential(1.0 / rate, size=n)))


def percentiles(latencies) -> dict:
    a = np.sort(np.asarray(list(latencies), dtype=float))
    return {"p50": float(np.percentile(a, 50)), "p99": float(np.percentile(a, 99)), "mean": float(a.mean())}


def main(argv=None) -> int:
    """python -m juliatdgl_amd.sweep [--mode streams|kernels] [--cycles K] [--workers W]: the
    64-run current sweep (sweep_specs) on the device, one line of runs/hour."""
    import argparse
    ap = argparse.ArgumentParser()
    ap.add_argument("--mode", choices=("streams", "kernels"), default="streams",
                    help="streams: concurrent handles (run_batch); kernels: one launch per kernel over all runs")
    ap.add_argument("--cycles", type=int, default=10)
    ap.add_argument("--workers", type=int, default=8)
    args = ap.parse_args(argv)
    specs = sweep_specs()
    work = Cycles(cycles=args.cycles)
    if args.mode == "streams":
        res, wall = run_batch(specs, work, workers=args.workers)
    else:
        res, wall = run_batched(specs, work, group=len(specs), workers=args.workers)
    print(f"{len(res)} runs of {args.cycles} cycles ({args.mode}): {wall:.3f} s, {3600.0 * len(res) / wall:.0f} runs/hour",
          flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
