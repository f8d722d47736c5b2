"""The ladder hook on the device (juliatdgl_amd.tdglgpu, the mirror of TDGLGPU's _gpu_cycle!,
gpu/TDGLGPU/src/TDGLGPU.jl:2114-2144): the policy and counters of test_backend_hook.jl with the HIP
stepper as the backend, and the Stepper-side state it writes back after every cycle as the CPU step
would leave it: the PCG history, the open boundary's clock and ramp (restarted per polarity as
run_ladders does, protocol.jl:852), the normal-fluid potentials and the noise stream's position.
Needs a HIP device."""
import numpy as np
import pytest

from cases import compare_states, mgb2

from juliatdgl_amd import tdglgpu
from juliatdgl_amd.scheme import Grid, box_masks, make_stepper, two_band, wedge_masks
from juliatdgl_amd.state import initial_state

pytestmark = pytest.mark.gpu


def _hip():
    from juliatdgl_amd.hip import device_count
    if device_count() < 1:
        pytest.fail("no HIP device visible")


def _core(errs):
    return {k: v for k, v in errs.items() if not k.startswith(("U", "V"))}


def test_backend_hook_policy_on_the_device():
    """test_backend_hook.jl:3-31 with the HIP stepper: two hook steps equal two CPU steps, the cache
    is held, re-enabling drops it and the counters, disabling keeps them."""
    from oracle.oracle import Oracle
    _hip()
    md = two_band(xi_ratio=4.0, kappa=2.0)
    gr = Grid.make(12, 8, 4, dx=0.5)
    st = make_stepper(md, gr, box_masks(gr))
    s = initial_state(md, gr, st.masks, 0.03)
    o = Oracle(st)
    o.set_state(s)
    o.step(2, 0.03, 0.0)
    ref = o.get_state()
    print("[gpu] hook policy", flush=True)
    try:
        with pytest.raises(ValueError):
            tdglgpu.enable(fallback="invalid")
        tdglgpu.enable(fallback="error")
        tdglgpu.cycle_steps(s, st, 2, 0.03, 0.0)
        q = tdglgpu.execution_stats()
        assert q["backend"] == "HIP" and q["backend_steps"] == 2 and q["cpu_fallback_steps"] == 0
        assert np.max(np.abs(s.G[0] - ref.G[0])) < 1e-12
        assert tdglgpu._state["last"] is not None and tdglgpu._state["dev"] is not None
        tdglgpu.enable()
        assert tdglgpu._state["last"] is None and tdglgpu._state["dev"] is None
        assert tdglgpu.execution_stats()["backend_steps"] == 0
        tdglgpu.cycle_steps(s, st, 1, 0.03, 0.0)
        assert tdglgpu.execution_stats()["backend_steps"] == 1
    finally:
        tdglgpu.disable()
    assert tdglgpu.STEPS[0] is None and tdglgpu._state["dev"] is None
    assert tdglgpu.execution_stats()["backend_steps"] == 1


def _ladder_through_hook(st, Js, H, cycles, lmax):
    """Both polarities through the hook, as run_ladders drives them (a fresh Meissner state and the
    open boundary restarted per polarity), `cycles` cycles per current; the final states."""
    out = {}
    for label, sign in (("J+", 1.0), ("J-", -1.0)):
        tdglgpu.restart_open(st)
        s = initial_state(st.model, st.grid, st.masks, H)
        for J in Js:
            for _ in range(cycles):
                tdglgpu.cycle_steps(s, st, lmax, H, st.selffield(sign * J))
        out[label] = (s, st.open_state, st.pcg_state)
    return out


def _ladder_oracle(st, Js, H, cycles, lmax):
    from oracle.oracle import Oracle
    out = {}
    for label, sign in (("J+", 1.0), ("J-", -1.0)):
        o = Oracle(st)
        o.set_state(initial_state(st.model, st.grid, st.masks, H))
        for J in Js:
            for _ in range(cycles):
                o.step(lmax, H, st.selffield(sign * J))
        out[label] = (o.get_state(), o.open_state(), o.get_pcg())
    return out


def test_open_boundary_ladder_through_the_hook_equals_the_oracle():
    """An open-boundary PCG run (KWT bands, buf 2) through the hook: two polarities of three
    currents, three cycles of 40 steps each, with h_δ recomputed every 7 steps (not a divisor of
    the cycle).  Each polarity restarts the boundary's ramp; the states, the ramps and the PCG
    histories the hook leaves in the Stepper equal the oracle's fresh run per polarity."""
    _hip()
    md = mgb2(gkwt=10.0)
    gr = Grid.make(24, 12, 6, dx=0.5, buf=2)
    kw = dict(sigma_vac=0.01, pcg_tol=1e-12, selffield_bc="open")
    dt = make_stepper(md, gr, wedge_masks(gr), **kw).dt
    st = make_stepper(md, gr, wedge_masks(gr), open_every=7 * dt, **kw)
    st_o = make_stepper(md, gr, wedge_masks(gr), open_every=7 * dt, **kw)
    assert st.open.every == 7
    Js, H, cyc, lmax = (0.0, 0.3, 0.6), 0.03, 3, 40
    print("[gpu] open-boundary ladder through the hook", flush=True)
    try:
        tdglgpu.enable(fallback="error")
        got = _ladder_through_hook(st, Js, H, cyc, lmax)
    finally:
        tdglgpu.disable()
    want = _ladder_oracle(st_o, Js, H, cyc, lmax)
    for label in ("J+", "J-"):
        (s, (cnt, old, new), pcg), (r, (ocnt, oold, onew), opcg) = got[label], want[label]
        errs = _core(compare_states(s, r))
        print(f"    {label}: worst {max(errs.values()):.2e}; open clock {cnt} / {ocnt}", flush=True)
        assert max(errs.values()) <= 1e-10, errs
        assert cnt == ocnt
        for a, b in zip(new, onew):
            assert np.max(np.abs(a - b)) <= 1e-10 * max(1.0, np.max(np.abs(b)))
        assert pcg.warm == opcg.warm and pcg.have2 == opcg.have2
        for a, b in zip(pcg.x, opcg.x):
            assert np.max(np.abs(a - b)) <= 1e-9 * max(1e-12, np.max(np.abs(b)))


def test_hook_writes_back_the_normal_fluid_potentials():
    """Two normal fluids and N normal fluids: after a hook cycle the Stepper holds δ (or μ_µ) of the
    last step as the oracle computes them (the conversion and Helmholtz solve of qp_potential!)."""
    from cases import qp_rates, three_frustrated
    from oracle.oracle import Oracle
    _hip()
    gr = Grid.make(24, 12, 6, dx=0.5)
    for md, qp in ((mgb2(), 1.0), (three_frustrated(), None)):
        nu = qp if qp is not None else qp_rates(0.5)(md)
        st = make_stepper(md, gr, wedge_masks(gr), qp_exchange=nu)
        s = initial_state(md, gr, st.masks, 0.03)
        s.G[0][...] *= np.exp(0.3j * np.linspace(0, 1, s.G[0].size)).reshape(s.G[0].shape, order="F")
        s.F[0][...] = s.G[0]
        o = Oracle(st)
        o.set_state(s)
        print(f"[gpu] normal-fluid potentials through the hook, {st.N} bands", flush=True)
        try:
            tdglgpu.enable(fallback="error")
            tdglgpu.cycle_steps(s, st, 5, 0.03, st.selffield(0.4))
        finally:
            tdglgpu.disable()
        o.step(4, 0.03, st.selffield(0.4))
        modes, _ = o.qp_potential()   # the potentials of the fifth step, from the state it starts at
        q = st.qp
        j0, k0 = q.j0, q.k0
        sl = (slice(1, gr.Nx), slice(j0, j0 + q.nj), slice(k0, k0 + q.nk))
        want = []
        if q.mode == 1:
            d = np.zeros_like(st.qp_state[0])
            d[sl] = modes[0]
            want.append(d)
        else:
            for mu in range(st.N):
                acc = np.zeros_like(modes[0])
                for kk in range(q.K):
                    acc = acc + q.P[mu, kk] * modes[kk]
                d = np.zeros_like(st.qp_state[0])
                d[sl] = acc
                want.append(d)
        assert len(st.qp_state) == len(want)
        scale = max(float(np.max(np.abs(w))) for w in want)
        assert scale > 0
        for a, b in zip(st.qp_state, want):
            assert np.max(np.abs(a - b)) <= 1e-11 * scale


def test_hook_keeps_the_noise_stream_position():
    """A noisy stepper through the hook: the Stepper's noise position advances with the steps, and
    putting it back replays the same draws (the state after the cycle is the same to the bit)."""
    _hip()
    md = mgb2()
    gr = Grid.make(24, 12, 6, dx=0.5)
    st = make_stepper(md, gr, wedge_masks(gr), noise=1e-3, noise_seed=11)
    s0 = initial_state(md, gr, st.masks, 0.03)
    print("[gpu] noise position through the hook", flush=True)
    try:
        tdglgpu.enable(fallback="error")
        a = s0.copy()
        tdglgpu.cycle_steps(a, st, 20, 0.03, 0.0)
        assert st.noise_counter == 20
        b = s0.copy()
        st.noise_counter = 0
        tdglgpu.cycle_steps(b, st, 20, 0.03, 0.0)
        assert st.noise_counter == 20
        c = s0.copy()
        tdglgpu.cycle_steps(c, st, 20, 0.03, 0.0)   # the stream goes on: other draws
    finally:
        tdglgpu.disable()
    assert all(np.array_equal(x, y) for x, y in zip(a.G, b.G))
    assert not np.array_equal(a.G[0], c.G[0])
