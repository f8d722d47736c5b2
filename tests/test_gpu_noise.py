"""Thermal noise on the device (noise.jl and the scheme it changes, bands.jl) against the oracle on
the same draws, with the generator on the device and with the caller's draws, and against the
reference's statistics (gtdgl/test/test_physics.jl:475-535, 697-733) on the device itself.  Needs
a HIP device."""
import numpy as np
import pytest

from juliatdgl_amd.noise import pack
from juliatdgl_amd.checkpoint import restore, snapshot
from cases import compare_states, make_case, seed_state
from test_oracle_noise import decoupled_equilibrium, field_equipartition, kwt_equilibrium_and_rates

pytestmark = pytest.mark.gpu

CASES = {
    # plain bands on RKC stages: the kick after the step
    "production": dict(noise=1e-3),
    # KWT on RKC stages: SK-ROCK, log-polar stages, KWT rotations, the low-passed frame
    "kwt10": dict(noise=1e-3),
    # KWT on Euler substeps: the polar kick after the step
    "kwt10_euler": dict(noise=1e-3),
    # KWT without the frame
    "kwt10_noframe": dict(noise=1e-3, rotframe=False),
    # the field kick with D (insulating padding, PCG) and SK-ROCK
    "kwt10_pcg": dict(noise=1e-3, noise_field=True),
    # the field kick without D (explicit field), hard nodes skipped
    "hard": dict(noise=1e-3, noise_field=True, implicit=False),
    # the normal-fluid turn before the kick
    "qp_two": dict(noise=1e-3),
}


def _hip():
    from juliatdgl_amd.hip import HipStepper, device_count
    if device_count() < 1:
        pytest.fail("no HIP device visible")
    return HipStepper


def _case(name, **kw):
    base = "kwt10" if name == "kwt10_noframe" else name
    extra = dict(CASES[name])
    extra.update(kw)
    import warnings
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        return make_case(base, **extra)


@pytest.mark.parametrize("name", list(CASES))
def test_noise_step_matches_the_oracle(name):
    """Twelve noisy steps in three calls (graph replays advance the stream) on the device and on
    the oracle from the same state: the same run."""
    from oracle.oracle import Oracle
    HipStepper = _hip()
    st, tol = _case(name)
    s0, pcg, H, HJ = seed_state(st)
    o = Oracle(st)
    o.set_state(s0)
    if pcg is not None:
        o.set_pcg(pcg)
    o.step(12, H, HJ)
    ref = o.get_state()
    print(f"[gpu] noise parity, {name}", flush=True)
    dev = HipStepper(st)
    dev.upload(s0)
    if pcg is not None:
        dev.set_pcg(pcg)
    for _ in range(3):
        dev.step(4, H, HJ)
    assert dev.noise_counter == o.noise_counter == 12
    errs = compare_states(dev.download(), ref)
    core = {k: v for k, v in errs.items() if not k.startswith(("U", "V"))}
    print(f"    worst {max(core, key=core.get)} {max(core.values()):.3e}", flush=True)
    assert max(core.values()) <= tol, errs
    # the noise moved the state: the noiseless run differs by far more than the tolerance
    st0, _ = make_case("kwt10" if name == "kwt10_noframe" else name,
                       **({"rotframe": False} if name == "kwt10_noframe" else {}),
                       **({"implicit": False} if name == "hard" else {}))
    o0 = Oracle(st0)
    o0.set_state(s0)
    if pcg is not None:
        o0.set_pcg(pcg)
    o0.step(12, H, HJ)
    assert np.max(np.abs(o0.get_state().G[0] - ref.G[0])) > 1e3 * tol


def test_host_draws_match_the_oracle():
    """The caller's draws (any normals, here numpy's) taken by the device and by the oracle: the
    same run; a step past them is refused."""
    from oracle.oracle import Oracle
    from juliatdgl_amd.hip import TDGLError
    from juliatdgl_amd.noise import from_stream, per_step
    HipStepper = _hip()
    st_d, tol = _case("kwt10_pcg")
    s0, pcg, H, HJ = seed_state(st_d)
    st, _ = _case("kwt10_pcg", noise_draws="host")
    rng = np.random.default_rng(21)
    flat = pack(st, [from_stream(st, rng.standard_normal(per_step(st))) for _ in range(8)])
    o = Oracle(st)
    o.set_state(s0)
    o.set_pcg(pcg)
    o.noise_draws(flat)
    o.step(8, H, HJ)
    print("[gpu] noise with the caller's draws", flush=True)
    dev = HipStepper(st)
    dev.upload(s0)
    dev.set_pcg(pcg)
    dev.noise_draws(flat)
    dev.step(5, H, HJ)
    dev.step(3, H, HJ)
    with pytest.raises(TDGLError, match="draws cover 0 more steps"):
        dev.step(1, H, HJ)
    errs = compare_states(dev.download(), o.get_state())
    core = {k: v for k, v in errs.items() if not k.startswith(("U", "V"))}
    print(f"    worst {max(core, key=core.get)} {max(core.values()):.3e}", flush=True)
    assert max(core.values()) <= tol, errs


def test_noise_snapshot_continues_to_the_bit():
    """A device snapshot holds the stream's position: continuing from it draws the same numbers."""
    HipStepper = _hip()
    st, _ = _case("kwt10_pcg")
    s0, pcg, H, HJ = seed_state(st)
    print("[gpu] noise snapshot", flush=True)
    a = HipStepper(st)
    a.upload(s0)
    a.set_pcg(pcg)
    a.step(6, H, HJ)
    sn = snapshot(a)
    assert sn.noise["noise_counter"] == 6
    a.step(6, H, HJ)
    b = HipStepper(st)
    restore(b, sn)
    b.step(6, H, HJ)
    sa, sb = a.download(), b.download()
    for x, y in zip(sa.G, sb.G):
        assert np.array_equal(x, y)
    assert np.array_equal(sa.hz, sb.hz)
    # another position of the stream, another run
    c = HipStepper(st)
    restore(c, sn)
    c.noise_counter = 1000
    c.step(6, H, HJ)
    assert not np.array_equal(c.download().G[0], sa.G[0])


def _dev_backend(st, s):
    HipStepper = _hip()
    d = HipStepper(st)
    d.upload(s)
    return d


def _dev_steps(d, field=False):
    def run(n):
        if n > 0:
            d.step(n, 0.0, 0.0)
    if field:
        def get_h():
            s = d.download()
            return (s.hx, s.hy, s.hz)
        return run, get_h
    return run, lambda: d.download().G[0]


def test_device_equilibrium_of_decoupled_nodes():
    print("[gpu] decoupled nodes", flush=True)
    decoupled_equilibrium(_dev_backend, _dev_steps)


def test_device_kwt_noise_equilibrium_and_rates():
    print("[gpu] KWT noise equilibrium", flush=True)
    kwt_equilibrium_and_rates(_dev_backend, _dev_steps)


@pytest.mark.parametrize("kappa,svac,buf,teq,tav", [(0.25, None, 1, 500.0, 3000.0), (1.0, None, 1, 100.0, 1000.0),
                                                    (0.25, 0.01, 2, 500.0, 3000.0)])
def test_device_field_noise_equipartition(kappa, svac, buf, teq, tav):
    print(f"[gpu] field equipartition kappa={kappa} svac={svac}", flush=True)
    r = field_equipartition(_dev_backend, _dev_steps, kappa, svac, buf, teq, tav)
    print(f"    ratio {r:.4f}", flush=True)
    assert r == pytest.approx(1.0, abs=0.01)


def test_device_kwt_noise_extended_system_at_the_production_step():
    """test_physics.jl:544-562: coupled nodes (m = 1, κ = 25) at the production step, where a KWT
    band takes RKC stages and the noise goes in with them (SK-ROCK, log-polar stages): the mean
    |Ψ|² does not depend on the mobility, with the rotating frame (the default) and without."""
    from juliatdgl_amd.scheme import Band, Grid, Model, box_masks, make_stepper
    from juliatdgl_amd.state import initial_state

    def media(gk, **kw):
        md = Model.make((Band.make(m=1.0, Gamma=2.0, gkwt=gk),), np.zeros((1, 1)), kappa=25.0)
        gr = Grid.make(10, 10, 2, dx=0.5)
        mk = box_masks(gr)
        st = make_stepper(md, gr, mk, noise=0.01, noise_seed=7, **kw)
        d = _dev_backend(st, initial_state(md, gr, mk, 0.0))
        d.step(round(40 / st.dt), 0.0, 0.0)
        sl = (slice(gr.x1, gr.x2), slice(gr.y1, gr.y2), slice(gr.z1, gr.z2))
        sm, N = 0.0, 0
        for _ in range(round(1500 / st.dt) // 13):
            d.step(13, 0.0, 0.0)
            x = np.abs(d.download().G[0][sl]) ** 2
            sm += x.sum()
            N += x.size
        return sm / N, st

    print("[gpu] KWT noise, extended system", flush=True)
    n0, _ = media(0.0)
    n3, st3 = media(3.0)
    nf, _ = media(3.0, rotframe=False)
    print(f"    <|psi|^2> {n0:.4f} (no KWT), {n3:.4f} (KWT), {nf:.4f} (KWT, no frame); stages {st3.rkl}", flush=True)
    assert st3.rotframe and any(st3.sk)
    assert n3 == pytest.approx(n0, rel=0.02)
    assert nf == pytest.approx(n0, rel=0.02)


def test_device_kwt_noise_on_the_notched_two_band_box():
    """test_physics.jl:566-582: the two-band wedge box at κ = 2 with γ̃ = 30 and noise stays finite on
    RKC stages (5, 13) and keeps ⟨|Ψπ|²⟩ within 5 % of γ̃ = 0."""
    from juliatdgl_amd.scheme import Grid, make_stepper, two_band, wedge_masks
    from juliatdgl_amd.state import initial_state
    gr = Grid.make(24, 12, 6, dx=0.5)
    mk = wedge_masks(gr)

    def media(gk):
        md = two_band(xi_ratio=4.0, kappa=2.0, gamma_kwt=gk)
        st = make_stepper(md, gr, mk, noise=1e-3, noise_seed=3)
        d = _dev_backend(st, initial_state(md, gr, mk, 0.03))
        HJ = st.selffield(0.0)
        sl = (slice(gr.x1, gr.x2), slice(gr.y1, gr.y2), slice(gr.z1, gr.z2))
        d.step(500, 0.03, HJ)
        a, c = 0.0, 0
        for _ in range(100):
            d.step(10, 0.03, HJ)
            a += np.sum(np.abs(d.download().G[1][sl]) ** 2)
            c += 1
        s = d.download()
        return all(np.all(np.isfinite(g)) for g in s.G), a / c, st

    print("[gpu] KWT noise, notched two-band box", flush=True)
    ok0, p0, _ = media(0.0)
    ok, p, st = media(30.0)
    print(f"    <|psi_pi|^2> sum {p0:.3f} (no KWT), {p:.3f} (γ̃ = 30); stages {st.rkl}", flush=True)
    assert ok0 and ok and st.rkl == (5, 13)
    assert p == pytest.approx(p0, rel=0.05)
