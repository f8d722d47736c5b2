"""TDGLGPU's interface over the HIP stepper (gpu/TDGLGPU/src/TDGLGPU.jl:29 exports).

    gs = GPUStepper(st)                 # st: a Stepper the HIP stepper runs (supported(st))
    g  = device_state(gs, s)            # s: a host State
    gpu_steps(g, gs, n, H=..., HJ=...)  # n steps on the device
    host_state(s, g)                    # back into s

and the ladder hook: enable(fallback=...) puts a cycle function into STEPS (the process-global
plugin point TDGL.STEPS, gtdgl/src/protocol.jl:301-316), disable() removes it, execution_stats()
counts backend and CPU-fallback cycles.  The CPU step is the caller's (TDGL.step! on the Julia
side): `fallback="cpu"` needs it passed as `cpu_step`; this package has no CPU step of its own.
"""
from __future__ import annotations

import warnings
from dataclasses import dataclass
from typing import Callable, Optional

from .hip import HipStepper, TDGLError, supported_desc
from .scheme import Stepper
from .state import State

__all__ = ["supported", "GPUStepper", "GState", "device_state", "device_state_", "host_state", "gpu_step",
           "gpu_steps", "enable", "disable", "execution_stats", "STEPS", "cycle_steps", "restart_open"]


def supported(st: Stepper):
    """supported(st) -> (ok, why) (TDGLGPU.jl:39-44)."""
    return supported_desc(st)


class GPUStepper:
    """GPUStepper(st, backend) (TDGLGPU.jl:1299-1316): errors when `st` has an option this
    version does not run."""

    def __init__(self, st: Stepper):
        ok, why = supported(st)
        if not ok:
            raise TDGLError(f"the HIP stepper does not run {why} yet (the CPU step does)")
        self.st = st
        self.dev = HipStepper(st)


@dataclass
class GState:
    """The device state: held by the stepper's handle (one run per handle)."""
    gs: GPUStepper
    template: State


def device_state(gs: GPUStepper, s: State) -> GState:
    """device_state(backend, s) (TDGLGPU.jl:64-66): the state copied to the device."""
    gs.dev.upload(s)
    return GState(gs, s.copy())


def device_state_(g: GState, s: State) -> GState:
    """device_state!(g, s) (TDGLGPU.jl:69-74)."""
    g.gs.dev.upload(s)
    return g


def host_state(s: State, g: GState) -> State:
    """host_state!(s, g): the device state back into s (every field, in place)."""
    return g.gs.dev.download(into=s)


def gpu_step(g: GState, gs: GPUStepper, *, H: float, HJ: float) -> GState:
    gs.dev.step(1, H, HJ)
    return g


def gpu_steps(g: GState, gs: GPUStepper, n: int, *, H: float, HJ: float) -> GState:
    """gpu_steps!(g, gs, n; H, HJ) (TDGLGPU.jl:2065-2071): n steps, synchronous."""
    gs.dev.step(n, H, HJ)
    return g


# ------------------------------------------------------------------------------ the ladder hook
STEPS: list = [None]          # TDGL.STEPS[]: a callable (s, st, n, H, HJ) -> s, or None for the CPU loop
_state = {"last": None, "dev": None, "fallback": "cpu", "cpu_step": None, "name": "", "bc": 0, "bs": 0, "cc": 0, "cs": 0}


def execution_stats() -> dict:
    """(backend, backend_cycles, backend_steps, cpu_fallback_cycles, cpu_fallback_steps)
    (TDGLGPU.jl:2086-2087)."""
    return dict(backend=_state["name"], backend_cycles=_state["bc"], backend_steps=_state["bs"],
                cpu_fallback_cycles=_state["cc"], cpu_fallback_steps=_state["cs"])


def enable(fallback: str = "cpu", cpu_step: Optional[Callable] = None) -> None:
    """enable!(backend; fallback) (TDGLGPU.jl:2103-2111): every ladder cycle runs on the HIP
    stepper; caches and counters start fresh.  An unsupported Stepper runs `cpu_step` with a
    warning (fallback="cpu") or raises (fallback="error")."""
    if fallback not in ("cpu", "error"):
        raise ValueError("fallback must be 'cpu' or 'error'")
    _state.update(last=None, dev=None, fallback=fallback, cpu_step=cpu_step, name="HIP", bc=0, bs=0, cc=0, cs=0)
    STEPS[0] = lambda s, st, n, H, HJ: _cycle(s, st, n, H, HJ)


def disable() -> None:
    """disable!() (TDGLGPU.jl:2112): the CPU loop again; the counters stay."""
    STEPS[0] = None
    _state.update(last=None, dev=None)


def _unsupported(why: str) -> None:
    if _state["fallback"] == "error":
        raise ValueError(f"HIP stepper: {why} is unsupported; CPU fallback is disabled")
    warnings.warn(f"HIP stepper: {why} is unsupported; this stepper runs on the CPU")


def _cycle(s: State, st: Stepper, n: int, H: float, HJ: float) -> State:
    """_gpu_cycle! (TDGLGPU.jl:2114-2144): upload, n steps, download."""
    last = _state["last"]
    if last is None or last[0] is not st:
        ok, why = supported(st)
        if not ok:
            _unsupported(why)
        _state["last"] = (st, GPUStepper(st) if ok else None)
        _state["dev"] = None
    gs = _state["last"][1]
    if gs is None:
        cpu = _state["cpu_step"]
        if cpu is None:
            raise ValueError("no CPU step was given to enable(cpu_step=...)")
        for _ in range(n):
            cpu(s, st, H, HJ)
        _state["cc"] += 1
        _state["cs"] += n
        return s
    dev = _state["dev"]
    if dev is None or dev[0] is not s:
        _state["dev"] = (s, device_state(gs, s))
    else:
        device_state_(dev[1], s)
    g = _state["dev"][1]
    _stepper_up(gs.dev, st)
    gpu_steps(g, gs, n, H=H, HJ=HJ)
    _stepper_down(gs.dev, st)
    host_state(s, g)
    _state["bc"] += 1
    _state["bs"] += n
    return s


# The Stepper's own mutable state, which the CPU step keeps in the Stepper and the checkpoints and
# measurements read there (TDGLGPU.jl:2130-2141): here attributes of the Python Stepper, absent
# until a cycle has set them (the state of a fresh Stepper).
#   st.pcg_state   PcgState: x, xprev, warm, have2, iters (_pcg_up! / _pcg_down!)
#   st.open_state  (count, old, new): the open boundary's clock and ramp (_open_up!; st.open.count,
#                  old, new); run_ladders restarts a polarity with count = -1 (protocol.jl:852)
#   st.qp_state    the normal-fluid potentials of the last step (st.qpx.δ or st.qpx.μ)
#   st.noise_counter  the noise stream's position (with the seed, the generator's whole state: what
#                  noise_rng is to the CPU step)
def _stepper_up(dev, st: Stepper) -> None:
    if st.field_mode == 2 and getattr(st, "pcg_state", None) is not None:
        dev.set_pcg(st.pcg_state)
    if st.selffield_bc == "open" and getattr(st, "open_state", None) is not None:
        dev.set_open(*st.open_state)
    if st.noise is not None and getattr(st, "noise_counter", None) is not None:
        dev.noise_counter = st.noise_counter


def _stepper_down(dev, st: Stepper) -> None:
    if st.field_mode == 2:
        st.pcg_state = dev.get_pcg()
    if st.selffield_bc == "open":
        st.open_state = dev.get_open()
    if st.qp is not None:
        st.qp_state = dev.get_qp()
    if st.noise is not None:
        st.noise_counter = dev.noise_counter


def restart_open(st: Stepper) -> None:
    """What run_ladders does to the open boundary before each polarity (protocol.jl:852): the
    ramp starts afresh (count = -1) at the next cycle."""
    if st.selffield_bc == "open" and getattr(st, "open_state", None) is not None:
        _, old, new = st.open_state
        st.open_state = (-1, old, new)


def cycle_steps(s: State, st: Stepper, n: int, H: float, HJ: float, cpu_step: Optional[Callable] = None) -> State:
    """_cycle_steps! (protocol.jl:306-316): the hook if set, else n CPU steps."""
    f = STEPS[0]
    if f is not None:
        return f(s, st, n, H, HJ)
    if cpu_step is None:
        raise ValueError("no hook enabled and no CPU step given")
    for _ in range(n):
        cpu_step(s, st, H, HJ)
    return s
