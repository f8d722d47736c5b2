"""The per-cycle measurements of a ladder on the CPU: the oracle's restatement (oracle/meter.py)
against closed forms and the reference's own checks, the verdict helpers against the reference's
vectors (gtdgl/test/test_protocol.jl:47-58), and crossing_times and the midplane current through
the oracle backend against test_protocol.jl:153-163 and test_physics.jl:74-92."""
import math

import numpy as np
import pytest

from juliatdgl_amd.meter import Voltmeter, crossing_times, crossings_now, late_crossings, pendiente, silent_now
from juliatdgl_amd.scheme import Grid, box_masks, make_stepper, two_band, wedge_masks
from juliatdgl_amd.state import initial_state
from oracle.meter import OracleMeter, midplane_current, path_phase, theta_A, plane_sum, unwrap

TWO_PI = 2 * math.pi


def test_verdict_helpers_match_the_reference_vectors():
    """A slip at 30 % of the point and a quiet second half is not silent; the static phase change
    of a current step is (test_protocol.jl:47-58)."""
    m, dt_cyc = 500, 0.8
    T1 = [0.0] + [0.0 if k < 150 else -TWO_PI for k in range(1, m + 1)]
    tabs = 6 * TWO_PI  # fill(-2π, 3, 2), all lines complete
    assert crossings_now(T1, tabs, m, dt_cyc) < 0.5
    assert not silent_now(T1, tabs, m, dt_cyc, 0.5)
    T0 = [0.0] + [0.3 * (1 - math.exp(-k / 20)) for k in range(1, m + 1)]
    assert silent_now(T0, 6 * 0.3, m, dt_cyc, 0.5)
    # the slope and the late turn by hand
    assert pendiente([0, 1, 2, 3], [1, 3, 5, 7]) == pytest.approx(2.0)
    assert late_crossings(T1, m) == pytest.approx(TWO_PI * m / (m - 125) / TWO_PI)


def _small():
    md = two_band(xi_ratio=4.0, kappa=2.0)
    gr = Grid.make(16, 8, 4, dx=0.5)
    mk = box_masks(gr)
    st = make_stepper(md, gr, mk)
    return md, gr, mk, st


def test_path_phase_is_gauge_invariant_and_counts_a_winding():
    """(conj(G_i) V_i) G_{i+1} is 1 for a pure gauge G = e^{iχ}, V = e^{-i(χ_{i+1} - χ_i)}; with
    V = 1 and G = e^{iqx} every link adds q."""
    md, gr, mk, st = _small()
    s = initial_state(md, gr, mk, 0.0)
    rng = np.random.default_rng(3)
    chi = rng.uniform(-3, 3, s.G[0].shape)
    s.G[0][...] = np.exp(1j * chi)
    s.Vx[...] = np.exp(-1j * (chi[1:] - chi[:-1]))  # Vx: one link fewer than nodes along x
    vi, vo = gr.x1 + 2, gr.x2 - 2
    D = path_phase(s.G[0], s, gr, vi, vo)
    assert np.max(np.abs(D)) < 1e-12
    q = 0.7
    x = np.arange(s.G[0].shape[0])[:, None, None]
    s.G[0][...] = np.exp(1j * q * x) * np.ones_like(chi)
    s.Vx[...] = 1.0
    D = path_phase(s.G[0], s, gr, vi, vo)
    assert np.allclose(D, (vo - vi) * q, rtol=0, atol=1e-12)


def test_unwrap_follows_the_phase_through_pi():
    T = np.zeros((2, 2))
    Dp = np.full((2, 2), 3.0)
    for step in range(1, 8):
        D = np.full((2, 2), (3.0 + 0.5 * step + math.pi) % TWO_PI - math.pi)
        unwrap(T, D, Dp)
        assert np.allclose(T, 0.5 * step)


def test_plane_sums_and_line_integral_by_hand():
    md, gr, mk, st = _small()
    s = initial_state(md, gr, mk, 0.0)
    x = np.arange(s.phx.shape[0])[:, None, None] + 1.0
    s.phx[...] = x * np.ones(s.phx.shape)
    vm = Voltmeter.of_ladder(st, 3)
    n = vm.Min.size
    assert plane_sum(s, gr, vm.vo, vm.Mou) / vm.Mou.sum() == pytest.approx(vm.vo)
    assert plane_sum(s, gr, vm.vi, vm.Min) == pytest.approx(vm.vi * n)
    assert theta_A(s, gr, vm.vi, vm.vo, vm.line_ok) == pytest.approx(sum(range(vm.vi, vm.vo)))


def test_midplane_current_of_a_linear_field():
    """hz = a·y gives jx = κ² a on every column of the mid plane: I = κ² a (Ny-1)(Nz-1) dy dz."""
    md, gr, mk, st = _small()
    s = initial_state(md, gr, mk, 0.0)
    a = 0.37
    y = np.arange(s.hz.shape[1])[None, :, None] * gr.dy
    s.hz[...] = a * y * np.ones(s.hz.shape)
    s.hy[...] = 0.0
    It, Is = midplane_current(s, st)
    assert It == pytest.approx(md.kappa ** 2 * a * (gr.Ny - 1) * (gr.Nz - 1) * gr.dy * gr.dz)
    ny_s = min(gr.y2 + 1, gr.Ny) - max(gr.y1, 2) + 1
    nz_s = min(gr.z2 + 1, gr.Nz) - max(gr.z1, 2) + 1
    assert Is == pytest.approx(md.kappa ** 2 * a * ny_s * nz_s * gr.dy * gr.dz)


def test_voltmeter_errors_as_the_reference():
    md, gr, mk, st = _small()
    with pytest.raises(ValueError, match="leaves no voltmeter"):
        Voltmeter.of_crossings(st, 20)
    vm = Voltmeter.of_ladder(st, 3)
    assert vm.line_ok.all() and vm.line_f1.all() and vm.Min.all()
    assert vm.Min.shape == (gr.y2 - gr.y1 + 2, gr.z2 - gr.z1 + 2)


def test_oracle_meter_first_cycle_and_settling():
    """The first cycle of a rung zeroes the unwrapped phases; a state that does not change has
    dif = diffh = 0 and keeps its phases."""
    from oracle.oracle import Oracle
    md, gr, mk, st = _small()
    s = initial_state(md, gr, mk, 0.03)
    o = Oracle(st)
    o.set_state(s)
    o.step(20, 0.03, st.selffield(0.2))
    s = o.get_state()
    vm = Voltmeter.of_ladder(st, 3)
    om = OracleMeter(st, vm, s)
    r1 = om.cycle(s, True)
    assert r1["T"] == [0.0, 0.0] and r1["dmax"] == 0.0 and r1["nbig"] == 0
    r2 = om.cycle(s, False)
    assert r2["dif"] == [0.0, 0.0] and r2["diffh"] == 0.0 and r2["T"] == [0.0, 0.0]
    assert r2["amax"] == pytest.approx(float(np.max(np.abs(s.F[0]))))


def test_crossing_times_on_the_oracle_matches_the_reference_test():
    """No crossing far below the onset, a regular train far above (test_protocol.jl:153-163)."""
    from oracle.oracle import Oracle
    from oracle.meter import OracleBackend
    md = two_band(xi_ratio=4.0, kappa=25.0)
    gr = Grid.make(24, 12, 4, dx=0.5)
    mk = wedge_masks(gr)
    st = make_stepper(md, gr, mk)
    o = Oracle(st)
    o.set_state(initial_state(md, gr, mk, 0.03))
    times, t, div = crossing_times(OracleBackend(o, st), st, 0.5, 0.03, 150.0, delta_v=6)
    assert times == [] and not div and abs(t - 150.0) < 0.5
    o = Oracle(st)
    o.set_state(initial_state(md, gr, mk, 0.03))
    times, t, div = crossing_times(OracleBackend(o, st), st, 1.4, 0.03, 600.0, delta_v=6, stop=4)
    assert len(times) == 4 and times == sorted(times) and t < 600.0 and not div


def test_midplane_current_carries_the_nominal_current():
    """With an insulating buffer and the bar self-field all the current goes through the sample and
    the total is the nominal one; with a normal-metal buffer part of it bypasses the sample
    (test_physics.jl:74-92)."""
    from oracle.oracle import Oracle
    md = two_band(xi_ratio=4.0, kappa=2.0)
    gr = Grid.make(24, 12, 4, dx=0.5, buf=4)
    mk = box_masks(gr)
    frac = {}
    for k, sv, bc in (("metal", None, "slab"), ("vacuum", 0.01, "bar")):
        kw = dict(selffield_bc=bc) if sv is None else dict(selffield_bc=bc, sigma_vac=sv)
        st = make_stepper(md, gr, mk, **kw)
        o = Oracle(st)
        o.set_state(initial_state(md, gr, mk, 0.03))
        o.step(round(30 / st.dt), 0.03, st.selffield(0.3))
        It, Is = midplane_current(o.get_state(), st)
        frac[k] = Is / It
        if bc == "bar":
            assert It / (-0.3 * 12 * 0.5 * 4 * 0.5) == pytest.approx(1.0, abs=2e-3)
    assert frac["metal"] < 0.95 and frac["vacuum"] > 0.995
