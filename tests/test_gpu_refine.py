"""The PCG with FP64 iterative refinement around Float32 inner solves (make_stepper(pcg_refine=True)):
against the oracle (the reference's solver: FP64 PCG with the Float32 preconditioner) to each case's
tolerance, eagerly and in captured cycles (with and without roll backs), and the error of the
three-band open-boundary workload against an all-FP64 run (FP64 PCG with the FP64 preconditioner),
stated for ψ and A.  Needs a HIP device."""
import numpy as np
import pytest

from cases import compare_states, make_case, seed_state

pytestmark = pytest.mark.gpu


def _hip():
    from juliatdgl_amd.hip import HipStepper, device_count
    if device_count() < 1:
        pytest.fail("no HIP device visible")
    return HipStepper


@pytest.mark.parametrize("graph", ["0", "1"])
@pytest.mark.parametrize("name", ["pcg_fast_f32", "pcg_tight_f32", "kwt10_pcg", "open_kwt_pcg", "qp_two_kwt_pcg",
                                  "pcg_mf1_f32"])
def test_refinement_matches_the_oracle(name, graph, monkeypatch):
    from oracle.oracle import Oracle
    HipStepper = _hip()
    monkeypatch.setenv("TDGL_HIP_GRAPH", graph)
    st, tol = make_case(name, pcg_refine=True)
    s0, pcg, H, HJ = seed_state(st)
    n = 20
    o = Oracle(st)
    o.set_state(s0)
    o.set_pcg(pcg)
    o.step(n, H, HJ)
    ref = o.get_state()
    print(f"[gpu] refinement {name}, graphs {graph}", flush=True)
    dev = HipStepper(st)
    dev.upload(s0)
    dev.set_pcg(pcg)
    dev.step(n, H, HJ)
    out = dev.download()
    stt = dev.stats()
    dev.close()
    errs = compare_states(out, ref)
    core = {k: v for k, v in errs.items() if not k.startswith(("U", "V"))}
    worst = max(core, key=core.get)
    print(f"    worst {worst} {core[worst]:.3e} (tol {tol:g}), {stt['pcg_iters']} inner iterations, "
          f"{stt['rollbacks']} roll backs", flush=True)
    assert max(core.values()) <= tol, errs
    assert stt["pcg_iters"] > 0


def test_refinement_rolls_back_a_short_budget(monkeypatch):
    """One captured round with a budget of one inner iteration cannot reach 1e-12: the cycle is
    rolled back and stepped eagerly, and the result still meets the oracle's."""
    from oracle.oracle import Oracle
    HipStepper = _hip()
    monkeypatch.setenv("TDGL_HIP_BUDGET", "1")
    st, tol = make_case("pcg_tight_f32", pcg_refine=True, pcg_rounds=1)
    s0, pcg, H, HJ = seed_state(st)
    o = Oracle(st)
    o.set_state(s0)
    o.set_pcg(pcg)
    o.step(12, H, HJ)
    dev = HipStepper(st)
    dev.upload(s0)
    dev.set_pcg(pcg)
    dev.step(12, H, HJ)
    stt = dev.stats()
    errs = compare_states(dev.download(), o.get_state())
    dev.close()
    print(f"[gpu] refinement roll back: {stt['rollbacks']} roll backs, worst {max(errs.values()):.2e}", flush=True)
    assert stt["rollbacks"] >= 1
    core = {k: v for k, v in errs.items() if not k.startswith(("U", "V"))}
    assert max(core.values()) <= tol, errs


def test_refinement_against_all_fp64_on_the_three_band_workload():
    """The three-band anisotropic open-boundary workload (buf 8, tol 1e-8) on 48x24x8 cells: the
    refinement against an all-FP64 run (FP64 preconditioner) after 80 steps; ψ and A agree to the
    solver's tolerance scale."""
    from juliatdgl_amd.configs import open_3band_256x128x32
    from juliatdgl_amd.state import seeded_state
    HipStepper = _hip()
    outs = []
    for kw in (dict(pcg_refine=True), dict(pcg_f32=False, pcg_fast=False)):
        st, p = open_3band_256x128x32(nx=48, ny=24, nz=8, **kw)
        H, HJ = p["H"], st.selffield(p["J"])
        s0 = seeded_state(st.model, st.grid, st.masks, H, seed=7)
        dev = HipStepper(st)
        dev.upload(s0)
        dev.step(80, H, HJ)
        outs.append(dev.download())
        dev.close()
    errs = compare_states(outs[0], outs[1])
    print("[gpu] refinement vs all-FP64, 48x24x8 three bands: " +
          ", ".join(f"{k} {errs[k]:.2e}" for k in ("G1", "G2", "G3", "thx", "thy", "thz")), flush=True)
    assert max(v for k, v in errs.items() if not k.startswith(("U", "V"))) <= 1e-6, errs


def test_refinement_onset_bins_against_all_fp64():
    """The golden ladder's sample with an insulating buffer (buf 2, σvac/σn = 0.01, PCG tol 1e-10):
    the two-polarity ladders stepped with the refinement have the verdicts, cycles and onset
    brackets of the same ladders stepped with every part of the PCG in FP64, and 400 steps at J = 0.6
    leave ψ and A within 1e-8 of them."""
    from juliatdgl_amd.meter import run_ladders
    from juliatdgl_amd.scheme import Grid, make_stepper, wedge_masks
    from juliatdgl_amd.state import initial_state
    from test_ladder import golden_setup
    HipStepper = _hip()
    _, _, st0, pr = golden_setup()
    gr = Grid.make(24, 12, 4, dx=0.5, buf=2)
    mk = wedge_masks(gr)
    kw = dict(selffield_bc="slab", rkl=False, sigma_vac=0.01, pcg_tol=1e-10)
    str_ = make_stepper(st0.model, gr, mk, pcg_refine=True, **kw)
    st64 = make_stepper(st0.model, gr, mk, pcg_fast=False, pcg_f32=False, **kw)
    print("[gpu] refinement against all-FP64: golden ladder sample", flush=True)
    out = []
    for st in (str_, st64):
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
    rr = run_ladders(lambda st: HipStepper(st), str_, Js, H=0.03, pr=pr)
    r64 = run_ladders(lambda st: HipStepper(st), st64, Js, H=0.03, pr=pr)
    for La, Lb in zip(rr[:2], r64[:2]):
        print(f"    {''.join(La.conv)} {list(La.m)} / {''.join(Lb.conv)} {list(Lb.m)}", flush=True)
        assert La.conv == Lb.conv and list(La.m) == list(Lb.m)
    print("    brackets " + ", ".join(f"{p} [{rr[2][p]['lo']}, {rr[2][p]['hi']}] / [{r64[2][p]['lo']}, {r64[2][p]['hi']}]"
                                   for p in ("pos", "neg")), flush=True)
    for pol in ("pos", "neg"):
        assert (rr[2][pol]["lo"], rr[2][pol]["hi"]) == (r64[2][pol]["lo"], r64[2][pol]["hi"])
