"""The sweep and the request loop (juliatdgl_amd.sweep) on the CPU: the seeds, the run
distribution, the footprint bound's monotonicity, and the admission of the server against its
budget and slots (with a stand-in for the device run)."""
import threading
import time

import pytest

from juliatdgl_amd import sweep
from juliatdgl_amd.sweep import RunSpec, Server, footprint, poisson_arrivals, percentiles, splitmix64, stepper_of, sweep_specs


def test_splitmix64_known_outputs():
    """The first outputs of SplitMix64 from state 0 (the published reference sequence)."""
    x, outs = 0, []
    for _ in range(3):
        x, v = splitmix64(x)
        outs.append(v)
    assert outs == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_sweep_specs_follow_the_distribution():
    a = sweep_specs(32, master=11)
    b = sweep_specs(32, master=11)
    assert a == b and len(a) == 64 and len({s.seed for s in a}) == 64
    for s in a:
        assert s.nx % 4 == 0 and 32 <= s.nx <= 160 and s.ny % 4 == 0 and 16 <= s.ny <= 96
        assert s.nz % 4 == 0 and 4 <= s.nz <= 24 and 0.1 <= s.depth <= 0.4
        assert 0.05 - 1e-12 <= abs(s.J) <= 0.4 + 1e-12
    assert sum(s.J > 0 for s in a) == 32 and sweep_specs(32, master=12) != a


def test_footprint_grows_with_the_grid():
    f = [footprint(stepper_of(RunSpec(0, n, 16, 4, 0.1))) for n in (32, 64, 128)]
    assert f[0] < f[1] < f[2]


def test_server_admits_within_budget_and_slots(monkeypatch):
    """Runs wait for a free slot and for room in the budget; every run is served once; the
    measured concurrency never exceeds the slots or the budget."""
    specs = [RunSpec(q, 32 + 8 * (q % 3), 16, 4, 0.1) for q in range(12)]
    need = {s.run_id: footprint(stepper_of(s)) for s in specs}
    live, peak = {"n": 0, "bytes": 0}, {"n": 0, "bytes": 0}
    lock = threading.Lock()

    def fake_execute(spec, work, st=None):
        with lock:
            live["n"] += 1
            live["bytes"] += need[spec.run_id]
            peak["n"] = max(peak["n"], live["n"])
            peak["bytes"] = max(peak["bytes"], live["bytes"])
        time.sleep(0.02)
        with lock:
            live["n"] -= 1
            live["bytes"] -= need[spec.run_id]
        return spec.run_id * 10, need[spec.run_id]

    monkeypatch.setattr(sweep, "_execute", fake_execute)
    budget = 2.5 * max(need.values())
    srv = Server(budget_bytes=budget, slots=3)
    for s, t in zip(specs, poisson_arrivals(len(specs), rate=200.0)):
        srv.submit(s, t)
    done = srv.run(work=None)
    assert sorted(done) == list(range(12)) and all(done[q]["result"] == 10 * q for q in done)
    assert peak["n"] <= 3 and peak["bytes"] <= budget
    assert all(d["latency"] >= 0 and d["finish"] >= d["start"] >= d["arrival"] - 1e-3 for d in done.values())
    p = percentiles(d["latency"] for d in done.values())
    assert p["p50"] <= p["p99"]
    with pytest.raises(ValueError, match="more than the budget"):
        s2 = Server(budget_bytes=10.0, slots=1)
        s2.submit(specs[0])
        s2.run(work=None)
