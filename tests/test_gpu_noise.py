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
