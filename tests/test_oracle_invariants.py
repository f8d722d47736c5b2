"""The oracle against the closed forms and invariants of the reference's physics tests
(gtdgl/test/test_physics.jl, gpu/TDGLGPU/test/test_invariants.jl): the parts of the step the
golden states do not reach (KWT, RKL/RKC stages, the rotating frame, N bands, anisotropy)."""
import math

import numpy as np
import pytest

from juliatdgl_amd.scheme import (Band, Grid, Model, RKCCoef, _rkc_ratio, box_masks, make_stepper, two_band,
                                  wedge_masks)
from juliatdgl_amd.state import initial_state
from oracle.oracle import Oracle, kwt_update, set_threads


def relax(st, s, n, H=0.0, HJ=0.0):
    o = Oracle(st)
    o.set_state(s)
    o.step(n, H, HJ)
    return o.get_state()


@pytest.mark.parametrize("psi,R,g", [(0.7 + 0.2j, 0.3 - 0.1j, 10.0), (1e-9 + 0j, 0.2 + 0.4j, 10.0),
                                     (0.4 - 0.9j, -0.5 + 0.05j, 2.0)])
def test_kwt_update_solves_the_2x2_system(psi, R, g):
    """test_physics.jl:51-59: w + γ²ψ Re(ψ*w) = √(1+γ²|ψ|²) R/Γ."""
    dt = 0.01
    w = (kwt_update(psi, dt * R, g * g) - psi) / dt
    q = 1 + g * g * abs(psi) ** 2
    M = np.array([[1 + g * g * psi.real ** 2, g * g * psi.real * psi.imag],
                  [g * g * psi.real * psi.imag, 1 + g * g * psi.imag ** 2]])
    x = np.linalg.solve(M, math.sqrt(q) * np.array([R.real, R.imag]))
    assert abs(w - complex(x[0], x[1])) <= 1e-12 * abs(complex(x[0], x[1]))


@pytest.mark.parametrize("u,g,rate", [(1.0, 0.0, 2.0), (5.79, 0.0, 0.3454), (5.79, 10.0, 0.03437)])
def test_amplitude_relaxation_rate(u, g, rate):
    """test_physics.jl:60-73: a uniform amplitude perturbation decays at 2ψ₀²/(u√(1+γ²ψ₀²))
    per unit time (the Euler factor 1 − λ dt per step)."""
    md = Model.make((Band.make(Gamma=u, gkwt=g),), [[0.0]], kappa=2.0)
    gr = Grid.make(8, 6, 4, dx=0.5)
    st = make_stepper(md, gr, box_masks(gr), rkl=False)
    s = initial_state(md, gr, st.masks, 0.0)
    s.G[0] *= (1 + 1e-6)
    s.F[0] = s.G[0].copy(order="F")
    i, j, k = 5, 4, 3
    a = relax(st, s, 1)
    d1 = abs(a.G[0][i, j, k]) - 1
    b = relax(st, a, 1)
    d2 = abs(b.G[0][i, j, k]) - 1
    assert abs((1 - d2 / d1) / st.dt - rate) <= 2e-3 * rate


def test_gamma_scaled_amplitude_relaxation():
    """test_physics.jl:7-24: one band at H = 0, the factor per explicit step is 1 − 2ψ₀²dt/Γ."""
    for G in (1.0, 3.0):
        md = Model.make((Band.make(Gamma=G),), [[0.0]], kappa=2.0)
        gr = Grid.make(8, 6, 4, dx=0.5)
        st = make_stepper(md, gr, box_masks(gr))
        s = initial_state(md, gr, st.masks, 0.0)
        s.G[0] *= (1 + 1e-6)
        s.F[0] = s.G[0].copy(order="F")
        a = relax(st, s, 1)
        b = relax(st, a, 1)
        d1 = abs(a.G[0][5, 4, 3]) - 1
        d2 = abs(b.G[0][5, 4, 3]) - 1
        assert abs(d2 / d1 - (1 - 2 * st.dt / G)) <= 1e-5


@pytest.mark.parametrize("s", range(2, 21))
def test_rkc_polynomial(s):
    """test_physics.jl:290-308: the stability polynomial of RKCCoef: the bridged ratio, |R| ≤ 1
    on [−2β, 0] and above 1 just past it, R = e^z + O(z²)."""
    k = RKCCoef.make(s)

    def R(z):
        a, b = 1.0, k.mu[0] * (1 + z / k.q)
        for j in range(1, s):
            a, b = b, k.mu[j] * b * (1 + z / k.q) + k.nu[j] * a
        return b

    b2 = _rkc_ratio(s)
    assert 0.761 * s * s <= b2 <= 0.781 * s * s and k.c[-1] == 1.0 and all(np.diff(k.c) > 0)
    assert 0.967 * s * s <= _rkc_ratio(s, 0.05) <= 0.971 * s * s
    zs = np.linspace(0, -2 * b2, 2001)
    assert max(abs(R(z)) for z in zs) <= 1 + 1e-12
    assert abs(R(-2.2 * b2)) > 1
    e1 = R(-1e-3) - math.exp(-1e-3)
    e2 = R(-5e-4) - math.exp(-5e-4)
    assert abs(e1) < 1e-6 and 3.9 < e1 / e2 < 4.1


def test_three_bands_trsb():
    """test_physics.jl:325-336: three identical bands with equal repulsive couplings relax to
    phases 0, ±2π/3 with |Ψ|² = 1 + |γ|."""
    g = -0.1
    md = Model.make((Band.make(), Band.make(), Band.make()), [[0, g, g], [g, 0, g], [g, g, 0]], kappa=25.0)
    gr = Grid.make(8, 8, 4, dx=0.5)
    mk = box_masks(gr)
    st = make_stepper(md, gr, mk)
    s = initial_state(md, gr, mk, 0.0)
    for mu, ph in ((1, 1.0), (2, 2.5)):
        s.G[mu] = s.G[mu] * np.exp(1j * ph)
        s.F[mu] = s.G[mu].copy(order="F")
    out = relax(st, s, round(200 / st.dt))
    c = ((gr.x1 + gr.x2 + 1) // 2 - 1, (gr.y1 + gr.y2 + 1) // 2 - 1, (gr.z1 + gr.z2 + 1) // 2 - 1)
    psi = [out.G[mu][c] for mu in range(3)]
    a2 = np.angle(psi[1] / psi[0])
    a3 = np.angle(psi[2] / psi[0])
    assert abs(abs(a2) - 2 * math.pi / 3) < 1e-6 and abs(abs(a3) - 2 * math.pi / 3) < 1e-6
    assert a2 * a3 < 0
    assert all(abs(abs(x) ** 2 - (1 + abs(g))) < 1e-6 for x in psi)


def test_identical_bands_stay_identical():
    """test_physics.jl:37-48: two identical bands (with RKL stages) stay bit-identical."""
    b = two_band(xi_ratio=4.0, kappa=25.0).bands
    md = Model.make((b[0], b[1], b[1]), [[0, 1e-3, 1e-3], [1e-3, 0, 0], [1e-3, 0, 0]], kappa=25.0)
    gr = Grid.make(24, 12, 4, dx=0.5)
    st = make_stepper(md, gr, wedge_masks(gr))
    assert st.rkl[0] == 0 and st.rkl[1] == st.rkl[2] and st.rkl[1] > 1
    s = initial_state(md, gr, st.masks, 0.03)
    out = relax(st, s, 400, 0.03, st.selffield(0.4))
    assert np.array_equal(out.G[1], out.G[2])
    assert np.abs(out.G[1]).max() > 0.1


@pytest.mark.slow
def test_rotating_frame_kwt_state_independent_of_dt():
    """test_physics.jl:259-288: with γ̃ = 30 and J = 0.6 the relaxed max|Ψ1| is the same (1e-5)
    at the production dt, at dt/4 and with Euler substeps; without the frame it is 0.05 higher;
    RKC stages give the same state."""
    set_threads(8)
    gr = Grid.make(24, 12, 4, dx=0.5)
    mk = wedge_masks(gr)
    md = two_band(xi_ratio=4.0, kappa=2.0, gamma_kwt=30.0)

    def run(**kw):
        st = make_stepper(md, gr, mk, **kw)
        s = initial_state(md, gr, mk, 0.03)
        out = relax(st, s, round(40 / st.dt), 0.03, st.selffield(0.6))
        return float(np.abs(out.G[0]).max())

    a = run(rotframe=True)
    b = run(rotframe=True, dt_scale=0.25)
    c = run(rotframe=True, rkl=False)
    assert abs(a - b) < 1e-5 and abs(a - c) < 1e-5 and a < 1.0
    assert run(rotframe=False) - a > 0.05
    assert abs(run(rkl="rkc") - a) < 1e-5
