"""The device meter (tdgl_meter_begin / tdgl_meter_cycle) against the oracle's restatement of a
ladder's per-cycle measurements (oracle/meter.py), and the host logic of juliatdgl_amd.meter on the
device against the reference's checks (gtdgl/test/test_protocol.jl:153-163).  Needs a HIP device."""
import math

import numpy as np
import pytest

from juliatdgl_amd.meter import Protocol, Voltmeter, crossing_times, run_rung
from juliatdgl_amd.scheme import Grid, make_stepper, two_band, wedge_masks
from juliatdgl_amd.state import initial_state
from cases import make_case, seed_state

pytestmark = pytest.mark.gpu

KEYS = ("V", "TA", "Tf1", "It", "Is", "amax", "diffh", "dmax")


def _hip():
    from juliatdgl_amd.hip import HipStepper, device_count
    if device_count() < 1:
        pytest.fail("no HIP device visible")
    return HipStepper


def _close(a, b, tol):
    for k in KEYS:
        assert a[k] == pytest.approx(b[k], rel=tol, abs=tol), (k, a[k], b[k])
    for k in ("T", "Tabs", "dif"):
        assert np.allclose(a[k], b[k], rtol=tol, atol=tol), (k, a[k], b[k])
    assert a["nbig"] == b["nbig"]


@pytest.mark.parametrize("name", ["production", "kwt10_pcg", "three_aniso"])
def test_device_meter_matches_the_oracle_meter_on_the_same_states(name):
    """Cycle by cycle, the record of the device equals the oracle's measurements of the state the
    device holds (downloaded after each cycle), to the rounding of the sums' order."""
    from oracle.meter import OracleMeter
    HipStepper = _hip()
    st, _ = make_case(name)
    s0, pcg, H, HJ = seed_state(st)
    vm = Voltmeter.of_ladder(st, 3)
    print(f"[gpu] meter on {name}", flush=True)
    dev = HipStepper(st)
    dev.upload(s0)
    if pcg is not None:
        dev.set_pcg(pcg)
    dev.meter_begin(vm)
    om = OracleMeter(st, vm, dev.download())
    for m in range(1, 6):
        dev.step(8, H, HJ)
        rd = dev.meter_cycle(m == 1)
        ro = om.cycle(dev.download(), m == 1)
        _close(rd, ro, 1e-12)
        assert math.isfinite(rd["amax"]) and rd["amax"] > 0.1
    assert abs(rd["It"]) > 0 and rd["diffh"] > 0


def test_device_meter_follows_the_oracle_run():
    """The device stepping and measuring against the oracle stepping and measuring: the records
    of the two runs agree to the parity of the step."""
    from oracle.meter import OracleBackend
    from oracle.oracle import Oracle
    HipStepper = _hip()
    st, tol = make_case("production")
    s0, pcg, H, HJ = seed_state(st)
    vm = Voltmeter.of_ladder(st, 3)
    o = Oracle(st)
    o.set_state(s0)
    ob = OracleBackend(o, st)
    dev = HipStepper(st)
    dev.upload(s0)
    for b in (ob, dev):
        b.meter_begin(vm)
    for m in range(1, 5):
        rs = []
        for b in (ob, dev):
            b.step(10, H, HJ)
            rs.append(b.meter_cycle(m == 1))
        _close(rs[1], rs[0], 1e-9)


def test_device_crossing_times_matches_the_reference_test():
    """No crossing far below the onset, a regular train far above (test_protocol.jl:153-163), with
    the measurements on the device every 10 steps."""
    HipStepper = _hip()
    md = two_band(xi_ratio=4.0, kappa=25.0)
    gr = Grid.make(24, 12, 4, dx=0.5)
    mk = wedge_masks(gr)
    st = make_stepper(md, gr, mk)
    print("[gpu] crossing_times below and above the onset", flush=True)
    dev = HipStepper(st)
    dev.upload(initial_state(md, gr, mk, 0.03))
    times, t, div = crossing_times(dev, st, 0.5, 0.03, 150.0, delta_v=6)
    assert times == [] and not div and abs(t - 150.0) < 0.5
    dev = HipStepper(st)
    dev.upload(initial_state(md, gr, mk, 0.03))
    times, t, div = crossing_times(dev, st, 1.4, 0.03, 600.0, delta_v=6, stop=4)
    assert len(times) == 4 and times == sorted(times) and t < 600.0 and not div


def test_device_rung_closes_as_the_oracle_rung():
    """One ladder point at a current below and one above the onset (the κ = 25 box): the device
    and the oracle close them with the same verdict after the same number of cycles."""
    from oracle.meter import OracleBackend
    from oracle.oracle import Oracle
    HipStepper = _hip()
    md = two_band(xi_ratio=4.0, kappa=25.0)
    gr = Grid.make(24, 12, 4, dx=0.5)
    mk = wedge_masks(gr)
    st = make_stepper(md, gr, mk)
    pr = Protocol(t_min_GL=20.0, t_min_GL_res=60.0, mMax=400, delta_v=6)
    vm = Voltmeter.of_ladder(st, pr.delta_v)
    for J in (0.5, 1.4):
        print(f"[gpu] rung at J = {J}", flush=True)
        o = Oracle(st)
        o.set_state(initial_state(md, gr, mk, 0.03))
        ob = OracleBackend(o, st)
        dev = HipStepper(st)
        dev.upload(initial_state(md, gr, mk, 0.03))
        out = []
        for b in (ob, dev):
            b.meter_begin(vm)
            out.append(run_rung(b, st, pr, J, 0.03))
        ro, rd = out
        print(f"    oracle {ro.conv} m={ro.m}, device {rd.conv} m={rd.m}", flush=True)
        assert rd.conv == ro.conv and rd.m == ro.m
        assert rd.conv == ("F" if J < 1 else "V")
        assert np.allclose(rd.V, ro.V, rtol=1e-8, atol=1e-8)


def test_golden_ladder_on_the_device():
    """The golden two-polarity ladder of release 1.2 (test_protocol.jl:4-29) with the device stepping
    and measuring: verdicts, cycles and brackets exactly, the recorded numbers and the final Ψ to 1e-10."""
    from test_ladder import check_golden_ladder
    HipStepper = _hip()
    print("[gpu] golden ladder", flush=True)
    check_golden_ladder(lambda st: HipStepper(st), report=lambda m: print(m, flush=True))


def test_fp32_preconditioner_against_all_fp64():
    """The golden ladder's sample with an insulating buffer (buf 2, σvac/σn = 0.01), stepped by the
    insulating PCG to 1e-10 with the FP32 preconditioner and FP64 outer iteration (the default) and
    with every part in FP64 (the exact path): 400 steps at J = 0.6 leave ψ and the links within
    1e-8 of each other, and the two-polarity ladders have the same verdicts, cycles and onset
    brackets."""
    from juliatdgl_amd.meter import run_ladders
    from test_ladder import golden_setup
    HipStepper = _hip()
    _, _, st0, pr = golden_setup()
    gr = Grid.make(24, 12, 4, dx=0.5, buf=2)
    mk = wedge_masks(gr)
    kw = dict(selffield_bc="slab", rkl=False, sigma_vac=0.01, pcg_tol=1e-10)
    st32 = make_stepper(st0.model, gr, mk, **kw)
    st64 = make_stepper(st0.model, gr, mk, pcg_fast=False, pcg_f32=False, **kw)
    assert st32.field_mode == st64.field_mode == 2 and st32.pcg_f32 and not st64.pcg_f32
    print("[gpu] FP32 preconditioner against all-FP64", flush=True)
    out = []
    for st in (st32, st64):
        d = HipStepper(st)
        d.upload(initial_state(st.model, st.grid, st.masks, 0.03))
        d.step(400, 0.03, st.selffield(0.6))
        out.append(d.download())
        d.close()
    a, b = out
    dpsi = max(float(np.max(np.abs(x - y))) for x, y in zip(a.G, b.G))
    dA = max(float(np.max(np.abs(getattr(a, k) - getattr(b, k)))) for k in ("phx", "phy", "phz"))
    print(f"    400 steps at J = 0.6: |Δψ| {dpsi:.2e}, |ΔA| {dA:.2e}", flush=True)
    assert dpsi <= 1e-8 and dA <= 1e-8
    Js = [0.30, 0.60, 0.90, 1.20, 1.50]
    r32 = run_ladders(lambda st: HipStepper(st), st32, Js, H=0.03, pr=pr)
    r64 = run_ladders(lambda st: HipStepper(st), st64, Js, H=0.03, pr=pr)
    for L32, L64 in zip(r32[:2], r64[:2]):
        print(f"    {''.join(L32.conv)} {list(L32.m)} / {''.join(L64.conv)} {list(L64.m)}", flush=True)
        assert L32.conv == L64.conv and list(L32.m) == list(L64.m)
    print("    brackets " + ", ".join(f"{p} [{r32[2][p]['lo']}, {r32[2][p]['hi']}] / [{r64[2][p]['lo']}, {r64[2][p]['hi']}]"
                                   for p in ("pos", "neg")), flush=True)
    for pol in ("pos", "neg"):
        assert (r32[2][pol]["lo"], r32[2][pol]["hi"]) == (r64[2][pol]["lo"], r64[2][pol]["hi"])
