"""The workloads of the measured configurations on the device, against the oracle and against each
other: the two-band KWT + insulating PCG run at its full 128x64x16 size; the three anisotropic
bands with an 8-cell insulating buffer, the open boundary and the Float32-preconditioned PCG, all
together on a reduced box and at the full 256x128x32 against the all-FP64 solve; a band-structure
plane wave through the device's anisotropic masses; and the served stream of noisy two- and
three-band requests (host-drawn normals: the same draws on the device and on the oracle).  Needs a
HIP device."""
import math

import numpy as np
import pytest

from cases import compare_states

from juliatdgl_amd import configs
from juliatdgl_amd.state import seeded_state

pytestmark = pytest.mark.gpu


def _hip():
    from juliatdgl_amd.hip import HipStepper, device_count
    if device_count() < 1:
        pytest.fail("no HIP device visible")
    return HipStepper


def _oracle_run(st, s0, n, H, HJ):
    from oracle.oracle import Oracle, set_threads
    set_threads(16)
    o = Oracle(st)
    o.set_state(s0)
    o.step(n, H, HJ)
    return o


def _core(errs):
    return {k: v for k, v in errs.items() if not k.startswith(("U", "V"))}


def test_kwt_pcg_128x64x16_full_size_against_the_oracle():
    """Two KWT bands (u = 5.79, γ̃ = 10) on RKC stages with the rotating frame, 128x64x16 cells, the
    insulating buffer and the PCG to 1e-10 (Float32 preconditioner), a 30-degree wedge: 8 steps
    from the seeded state on the device and on the oracle.  The PCG solutions agree to the
    tolerance times the conditioning, not to the bit: the bound is 1e-8."""
    HipStepper = _hip()
    st, p = configs.kwt_pcg_128x64x16()
    H, HJ = p["H"], st.selffield(p["J"])
    s0 = seeded_state(st.model, st.grid, st.masks, H, seed=20261202)
    n = 8
    print(f"[gpu] kwt_pcg_128x64x16: {n} steps, dt {st.dt:.5f}, stages {st.rkl}", flush=True)
    o = _oracle_run(st, s0, n, H, HJ)
    ref = o.get_state()
    dev = HipStepper(st)
    dev.upload(s0)
    dev.step(n, H, HJ)
    out = dev.download()
    errs = compare_states(out, ref)
    core = _core(errs)
    worst = max(core, key=core.get)
    gi, oi = dev.stats()["pcg_iters"], o.total_iters()
    print(f"    worst {worst} {core[worst]:.3e}; PCG iterations device {gi}, oracle {oi}", flush=True)
    assert max(core.values()) <= 1e-8, errs
    assert abs(gi - oi) <= max(4, oi // 4)
    dev.close()


@pytest.mark.parametrize("mode", ["f32", "exact"])
def test_open_3band_reduced_against_the_oracle(mode):
    """Everything of the 256x128x32 workload at once on 48x24x8 cells (64x40x24 with the 8-cell
    buffer): three anisotropic bands (one on RKC stages), the insulating buffer, the open boundary
    (its first recomputation and ramp inside the run) and the PCG, with the Float32 preconditioner
    at the workload's tolerance 1e-8 (bound 1e-6) and all-FP64 at 1e-12 (bound 1e-10)."""
    HipStepper = _hip()
    if mode == "f32":
        st, p = configs.open_3band_256x128x32(nx=48, ny=24, nz=8)
        tol = 1e-6
    else:
        st, p = configs.open_3band_256x128x32(nx=48, ny=24, nz=8, pcg_tol=1e-12, pcg_fast=False, pcg_f32=False)
        tol = 1e-10
    assert st.N == 3 and st.field_mode == 2 and st.selffield_bc == "open"
    assert any(b.ma[2] == 4 * b.ma[0] for b in st.model.bands)
    H, HJ = p["H"], st.selffield(p["J"])
    s0 = seeded_state(st.model, st.grid, st.masks, H, seed=31)
    n = 8
    print(f"[gpu] open_3band 48x24x8 {mode}: {n} steps, stages {st.rkl}, open every {st.open.every}", flush=True)
    o = _oracle_run(st, s0, n, H, HJ)
    ref = o.get_state()
    dev = HipStepper(st)
    dev.upload(s0)
    dev.step(n, H, HJ)
    out = dev.download()
    errs = compare_states(out, ref)
    core = _core(errs)
    worst = max(core, key=core.get)
    print(f"    worst {worst} {core[worst]:.3e} (bound {tol:g})", flush=True)
    assert max(core.values()) <= tol, errs
    cnt, old, new = dev.get_open()
    ocnt, oold, onew = o.open_state()
    assert cnt == ocnt
    for a, b in zip(new, onew):
        assert np.max(np.abs(a - b)) <= max(tol, 1e-12) * max(1.0, np.max(np.abs(b)))
    dev.close()


def test_open_3band_256x128x32_fp32_against_all_fp64():
    """The full 256x128x32 workload against the same run with every part of the PCG in FP64 (the
    preconditioner too): with the Float32 preconditioner inside the FP64 PCG (the reference's solver)
    and with the FP64 iterative refinement around Float32 inner solves (pcg_refine), ψ and A after
    five ladder cycles (200 steps from the seeded state, five recomputations of the open boundary).
    The tolerance of every solve is the workload's 1e-8; the states agree to a few times it."""
    HipStepper = _hip()
    out = {}
    for name, kw in (("f64", dict(pcg_f32=False)), ("f32", {}), ("refine", dict(pcg_refine=True))):
        st, p = configs.open_3band_256x128x32(**kw)
        H, HJ = p["H"], st.selffield(p["J"])
        s0 = seeded_state(st.model, st.grid, st.masks, H, seed=20261202)
        print(f"[gpu] open_3band 256x128x32 {name}: 5 cycles", flush=True)
        dev = HipStepper(st)
        dev.upload(s0)
        for _ in range(5):
            dev.step(40, H, HJ)
        out[name] = (dev.download(), dev.stats())
        dev.close()
    b, sb = out["f64"]
    for name in ("f32", "refine"):
        a, sa = out[name]
        dpsi = max(float(np.max(np.abs(x - y)) / np.max(np.abs(y))) for x, y in zip(a.G, b.G))
        dA = max(float(np.max(np.abs(getattr(a, k) - getattr(b, k)))) for k in ("phx", "phy", "phz")) / \
            max(float(np.max(np.abs(getattr(b, k)))) for k in ("phx", "phy", "phz"))
        dh = max(float(np.max(np.abs(getattr(a, k) - getattr(b, k)))) for k in ("hx", "hy", "hz"))
        print(f"    {name} against all-FP64: |Δψ|/|ψ| {dpsi:.2e}, |ΔA|/|A| {dA:.2e}, |Δh| {dh:.2e}; PCG "
              f"iterations {sa['pcg_iters']} (all-FP64 {sb['pcg_iters']})", flush=True)
        assert all(np.isfinite(g).all() for g in a.G)
        assert dpsi <= 1e-6 and dA <= 1e-6


def test_anisotropic_plane_wave_on_the_device():
    """The plane-wave closed form of the anisotropic masses (one band, m_axes = (0.8, 0.8, 3.2)) on
    the device: the Euler factor along each axis is 1 − 2c − dt/(m_i Γ dx²)·λ."""
    from juliatdgl_amd.scheme import Band, Grid, Model, box_masks, make_stepper
    from juliatdgl_amd.state import initial_state
    HipStepper = _hip()
    ma = (0.8, 0.8, 3.2)
    md = Model.make((Band.make(a=1.0, m_axes=ma),), [[0.0]], kappa=2.0)
    gr = Grid.make(8, 8, 8, dx=0.5, buf=1)
    dt = 0.002
    st = make_stepper(md, gr, box_masks(gr), rkl=False, dt=dt)
    rate = {}
    for axis in range(3):
        s = initial_state(md, gr, st.masks, 0.0)
        lo, hi = (gr.x1, gr.y1, gr.z1)[axis], (gr.x2, gr.y2, gr.z2)[axis]
        nint = hi - lo
        prof = np.zeros(s.G[0].shape[axis])
        prof[lo:hi] = np.cos(math.pi * (np.arange(1, nint + 1) - 0.5) / nint)
        prof[lo - 1], prof[hi] = prof[lo], prof[hi - 1]
        shp = [1, 1, 1]
        shp[axis] = -1
        G = st.masks.tpsi * (1.0 + 1e-7 * prof.reshape(shp))
        s.G[0][...] = G
        s.F[0][...] = G
        dev = HipStepper(st)
        dev.upload(s)
        dev.step(1, 0.0, 0.0)
        out = dev.download().G[0]
        dev.close()
        sl = tuple(slice(a, b + 1) for a, b in ((gr.x1, gr.x2 - 1), (gr.y1, gr.y2 - 1), (gr.z1, gr.z2 - 1)))
        d0, d1 = (G[sl] - 1).real, (out[sl] - 1).real
        big = np.abs(d0) > 0.3e-7
        g = float((d1[big] / d0[big]).mean())
        lam = 2 - 2 * math.cos(math.pi / nint)
        want = dt / (ma[axis] * gr.dx ** 2) * lam
        rate[axis] = 1 - g - 2 * dt
        print(f"[gpu] plane wave along axis {axis}: excess {rate[axis]:.6e}, closed form {want:.6e}", flush=True)
        assert abs(rate[axis] - want) <= 1e-5 * want
    assert abs(rate[0] / rate[2] - 4.0) <= 1e-4


def test_served_noisy_stream_matches_alone_and_the_oracle():
    """Requests of the served sweep (two and three bands, notch or hole, noise T̃ = 1e-3 from each
    run's own host stream: the validation mode), on grids drawn log-uniform from 24x16x4 to
    40x24x8: served from a Poisson stream they give the records they give alone, to the bit, and
    each run's state after its cycles equals the oracle's on the same draws."""
    from oracle.meter import OracleBackend
    from oracle.oracle import Oracle
    from juliatdgl_amd.hip import HipStepper as _HS
    from juliatdgl_amd.sweep import (Cycles, HostDraws, Server, poisson_arrivals, run_batch, serving_specs, state_of,
                                     stepper_of)
    _hip()
    specs = serving_specs(5, master=99, lo=(24, 16, 4), hi=(40, 24, 8), noise_host=True)
    assert {s.bands for s in specs} == {2, 3}
    print("[gpu] served noisy stream: " + ", ".join(f"{s.nx}x{s.ny}x{s.nz}/{s.bands}b" for s in specs), flush=True)
    work = Cycles(cycles=2, lmax=20)
    alone, _ = run_batch(specs, work, workers=1)
    srv = Server(budget_bytes=200e9, slots=3)
    for s, t in zip(specs, poisson_arrivals(len(specs), rate=30.0, seed=3)):
        srv.submit(s, t)
    done = srv.run(work)
    for s in specs:
        assert done[s.run_id]["result"] == alone[s.run_id], s.run_id
    for s in specs[:3]:
        st = stepper_of(s)
        dev = HostDraws(_HS(st), st, s.seed)
        dev.upload(state_of(s, st))
        ob = HostDraws(OracleBackend(Oracle(st), st), st, s.seed)
        ob.o.set_state(state_of(s, st))
        for b in (dev, ob):
            b.step(40, s.H, st.selffield(s.J))
        errs = _core(compare_states(dev.download(), ob.o.get_state()))
        print(f"    run {s.run_id} ({st.N} bands): worst {max(errs.values()):.2e}", flush=True)
        assert max(errs.values()) <= 1e-10, errs
        dev.close()
