"""The ladder hook's policy, cache and counters on the CPU (gpu/TDGLGPU/test/test_backend_hook.jl:3-31):
the fallback setting is validated, `fallback="error"` turns an unsupported Stepper into an
error, re-enabling drops the cached stepper and device state and zeroes the counters, disabling
restores the CPU loop and keeps them, and an unsupported Stepper runs the caller's CPU step
(here the oracle's) and is counted as such."""
import warnings

import numpy as np
import pytest

from juliatdgl_amd import tdglgpu
from juliatdgl_amd.scheme import Grid, box_masks, make_stepper, two_band
from juliatdgl_amd.state import initial_state
from oracle.oracle import Oracle


def _setup():
    md = two_band(xi_ratio=4.0, kappa=2.0)
    gr = Grid.make(12, 8, 4, dx=0.5)
    st = make_stepper(md, gr, box_masks(gr))
    s = initial_state(md, gr, st.masks, 0.03)
    return st, s


def _oracle_step():
    """A CPU step (the caller's TDGL.step! on the Julia side) by the oracle, one per stepper."""
    cache = {}

    def step(s, st, H, HJ):
        o = cache.get(id(st))
        if o is None:
            o = cache[id(st)] = Oracle(st)
        o.set_state(s)
        o.step(1, H, HJ)
        out = o.get_state()
        for (_, a), (_, b) in zip(s.arrays(), out.arrays()):
            a[...] = b
    return step


def test_fallback_setting_is_validated():
    with pytest.raises(ValueError):
        tdglgpu.enable(fallback="invalid")


def test_policy_cache_and_counters(monkeypatch):
    st, s = _setup()
    ref = s.copy()
    cpu = _oracle_step()
    for _ in range(2):
        cpu(ref, st, 0.03, 0.0)
    try:
        tdglgpu.enable(fallback="error", cpu_step=cpu)
        with pytest.raises(ValueError, match="rejected configuration"):
            tdglgpu._unsupported("rejected configuration")
        # an unsupported Stepper with fallback="error" is an error at its first cycle
        monkeypatch.setattr(tdglgpu, "supported", lambda st: (False, "rejected configuration"))
        with pytest.raises(ValueError):
            tdglgpu.cycle_steps(s.copy(), st, 2, 0.03, 0.0)
        # re-enabling drops the cache and the counters
        tdglgpu.enable(cpu_step=cpu)
        assert tdglgpu._state["last"] is None and tdglgpu._state["dev"] is None
        assert tdglgpu.execution_stats()["backend_steps"] == 0
        with pytest.warns(UserWarning, match="rejected configuration"):
            tdglgpu._unsupported("rejected configuration")
        # the unsupported Stepper runs the CPU step, counted as fallback cycles
        with warnings.catch_warnings():
            warnings.simplefilter("ignore")
            tdglgpu.cycle_steps(s, st, 2, 0.03, 0.0)
        q = tdglgpu.execution_stats()
        assert q["backend"] == "HIP" and q["cpu_fallback_cycles"] == 1 and q["cpu_fallback_steps"] == 2
        assert q["backend_steps"] == 0
        assert np.max(np.abs(s.G[0] - ref.G[0])) < 1e-12
        assert tdglgpu._state["last"] is not None and tdglgpu._state["last"][1] is None
    finally:
        tdglgpu.disable()
    assert tdglgpu.STEPS[0] is None and tdglgpu._state["dev"] is None
    assert tdglgpu.execution_stats()["cpu_fallback_steps"] == 2


def test_cycle_without_hook_needs_a_cpu_step():
    st, s = _setup()
    tdglgpu.disable()
    with pytest.raises(ValueError):
        tdglgpu.cycle_steps(s, st, 1, 0.03, 0.0)
    tdglgpu.cycle_steps(s, st, 1, 0.03, 0.0, cpu_step=_oracle_step())


def test_restart_open_resets_the_clock_only():
    st, _ = _setup()
    st.selffield_bc = "open"
    old = [np.ones(3)]
    new = [np.full(3, 2.0)]
    st.open_state = (17, old, new)
    tdglgpu.restart_open(st)
    assert st.open_state[0] == -1 and st.open_state[1] is old and st.open_state[2] is new
