"""Several runs on the device at once (juliatdgl_amd.sweep): every run's result is the same alone
and in the mix, the server's footprint bound holds, and noisy runs served from a Poisson stream end
at the same first phase slip as alone.  Needs a HIP device."""
import pytest

from juliatdgl_amd import sweep
from juliatdgl_amd.sweep import Cycles, FirstSlip, RunSpec, Server, footprint, poisson_arrivals, run_batch, stepper_of

pytestmark = pytest.mark.gpu


def _specs():
    out = []
    for q, (nx, ny, nz) in enumerate(((32, 16, 4), (40, 24, 8), (64, 16, 4), (48, 32, 12), (36, 20, 8), (56, 24, 4))):
        out.append(RunSpec(q, nx, ny, nz, (0.1 + 0.05 * q) * (-1) ** q, depth=0.1 + 0.05 * q, seed=100 + q))
    return out


def test_runs_are_the_same_alone_and_in_the_mix():
    specs = _specs()
    print("[gpu] sweep: alone and in the mix", flush=True)
    alone, t1 = run_batch(specs, Cycles(cycles=6), workers=1)
    mix, t6 = run_batch(specs, Cycles(cycles=6), workers=6)
    print(f"    one at a time {t1:.2f} s, six at a time {t6:.2f} s", flush=True)
    for q in alone:
        assert alone[q] == mix[q], q


def test_served_runs_fit_their_footprint_and_match_alone():
    specs = _specs()[:4]
    print("[gpu] served runs", flush=True)
    srv = Server(budget_bytes=2.2 * max(footprint(stepper_of(s)) for s in specs), slots=3)
    for s, t in zip(specs, poisson_arrivals(len(specs), rate=20.0)):
        srv.submit(s, t)
    done = srv.run(Cycles(cycles=4))
    alone, _ = run_batch(specs, Cycles(cycles=4), workers=1)
    for s in specs:
        d = done[s.run_id]
        assert 0 < d["bytes"] <= footprint(stepper_of(s)), (d["bytes"], footprint(stepper_of(s)))
        assert d["result"] == alone[s.run_id]


def test_noisy_runs_end_at_the_same_first_slip_served():
    """Requests with thermal noise (each its own counter-based stream) that end at their first phase
    slip: the same slip time served in the mix as alone."""
    specs = [RunSpec(q, 40, 20, 4, 1.2 * (-1) ** q, depth=0.2, seed=7 + q, noise=1e-3) for q in range(4)]
    print("[gpu] noisy runs to the first slip", flush=True)
    work = FirstSlip(T=60.0, every=10)
    alone, _ = run_batch(specs, work, workers=1)
    srv = Server(budget_bytes=200e9, slots=4)
    for s in specs:
        srv.submit(s, 0.0)
    done = srv.run(work)
    print("    " + ", ".join(f"{q}: {alone[q][0]}" for q in sorted(alone)), flush=True)
    for q in alone:
        assert done[q]["result"] == alone[q]


@pytest.mark.parametrize("fold64", ["default", "1"])
def test_batched_kernels_match_the_runs_alone(fold64, monkeypatch):
    """tdgl_batch_*: six uneven runs advanced by one launch per kernel over all of them give the
    records of their own steppers to the bit (the batch launches the instances each run's own step
    takes), in a batch of one as in the batch of six; with the full-length FP64 solve these small
    runs take by default, and with the parity folds (TDGL_HIP_FOLD64=1: the batched fold, quarter
    planes, quarter z pass and unfold)."""
    from juliatdgl_amd.sweep import run_batched
    if fold64 != "default":
        monkeypatch.setenv("TDGL_HIP_FOLD64", fold64)
    specs = _specs()
    print(f"[gpu] batched kernels, FP64 folds {fold64}", flush=True)
    alone, _ = run_batch(specs, Cycles(cycles=3), workers=1)
    mix, t6 = run_batched(specs, Cycles(cycles=3), group=6)
    one, _ = run_batched(specs, Cycles(cycles=3), group=1)
    print(f"    six in one batch {t6:.2f} s", flush=True)
    for q in alone:
        assert mix[q] == one[q], q
        assert mix[q] == alone[q], q


def test_batched_kernels_match_the_oracle():
    """Two uneven runs stepped 40 times in one batch against the oracle's step on the same seeded
    states: ψ and the links within 1e-10."""
    import numpy as np
    from oracle.oracle import Oracle
    from juliatdgl_amd.hip import HipBatch, HipStepper
    from juliatdgl_amd.sweep import state_of
    specs = _specs()[:2]
    print("[gpu] batched kernels against the oracle", flush=True)
    sts = [stepper_of(s) for s in specs]
    devs = []
    for s, st in zip(specs, sts):
        d = HipStepper(st)
        d.upload(state_of(s, st))
        devs.append(d)
    b = HipBatch(devs)
    H = np.array([s.H for s in specs])
    HJ = np.array([st.selffield(s.J) for s, st in zip(specs, sts)])
    b.step(40, H, HJ)
    for s, st, d, hj in zip(specs, sts, devs, HJ):
        o = Oracle(st)
        o.set_state(state_of(s, st))
        o.step(40, s.H, hj)
        ref, got = o.get_state(), d.download()
        worst = 0.0
        for x, y in zip(ref.G, got.G):
            worst = max(worst, float(np.max(np.abs(x - y)) / max(1.0, np.max(np.abs(x)))))
        for k in ("phx", "phy", "phz", "hx", "hy", "hz"):
            x, y = getattr(ref, k), getattr(got, k)
            worst = max(worst, float(np.max(np.abs(x - y)) / max(1.0, np.max(np.abs(x)))))
        print(f"    run {s.run_id}: worst {worst:.3e}", flush=True)
        assert worst <= 1e-10
    b.close()
    for d in devs:
        d.close()
