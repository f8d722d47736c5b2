"""Stepper cases of the parity tests, after the reference's step comparison
(gpu/TDGLGPU/test/compare_cpu.jl:28-38, 44-63): the two-band MgB2-like model of
gtdgl/materials/mgb2_model_II.toml at κ = 2 on a 24x12x6 WEDGE box, with one scheme option
changed per case, plus three bands, KWT and anisotropic masses; and the comparison itself
(compare_cpu.jl:65-109): 5 CPU steps off the seed, then n steps of the oracle and of the HIP
stepper from the same state (and the same PCG history), every field compared, vector fields
against the size of the whole vector (:91-96)."""
from __future__ import annotations

import numpy as np

from juliatdgl_amd.scheme import Band, Grid, Model, box_masks, make_stepper, wedge_masks
from juliatdgl_amd.state import initial_state


def mgb2(kappa=2.0, gkwt=0.0, sigma_n=1.0):
    """mgb2_model_II.toml: σ (m=1, Γ=1, a=1), π (ξ ratio 4, a=0.59), γ12 = 1e-3."""
    R, a, beta = 4.0, 0.59, 1.0
    m = 1 / (R * R * (a / beta))
    return Model.make((Band.make(gkwt=gkwt), Band.make(m=m, Gamma=1.0, a=a, gkwt=gkwt)),
                      [[0, 1e-3], [1e-3, 0]], kappa=kappa, sigma_n=sigma_n)


def three_frustrated(kappa=2.0):
    g = -0.1
    return Model.make((Band.make(), Band.make(), Band.make()), [[0, g, g], [g, 0, g], [g, g, 0]], kappa=kappa)


def three_aniso(kappa=2.0):
    """Three bands with anisotropic masses m_z = 4 m_xy: a = 1, 0.6, 0.3, ξ = 1, 1.8, 3."""
    bands = []
    for a, xi in ((1.0, 1.0), (0.6, 1.8), (0.3, 3.0)):
        m = 1 / (xi * xi * a)
        bands.append(Band.make(a=a, m_axes=(m, m, 4 * m)))
    gm = [[0, -0.1, -0.05], [-0.1, 0, -0.02], [-0.05, -0.02, 0]]
    return Model.make(tuple(bands), gm, kappa=kappa)


def qp_rates(r):
    """Exchange rates with G_µν = r for every pair (detailed balance, test_physics.jl:404):
    ν_µν = r / χ_µ with χ_µ = s_µ σn Γ_µ / g_µ and the dirty shares s_µ ∝ g_µ Γ_µ / β_µ."""
    def rates(md):
        w = [b.stiffness() * b.Gamma / b.beta for b in md.bands]
        s = [x / sum(w) for x in w]
        chi = [s[m] * md.sigma_n * md.bands[m].Gamma / md.bands[m].stiffness() for m in range(md.N)]
        return [[0.0 if m == n else r / chi[m] for n in range(md.N)] for m in range(md.N)]
    return rates


SMALL = dict(nx=24, ny=12, nz=6, dx=0.5)

# name -> (model factory, mesh overrides, scheme keywords, tolerance)
CASES = {
    "production": (mgb2, {}, {}, 1e-10),
    "euler_substeps": (mgb2, {}, dict(rkl=False), 1e-10),
    "no_extrap": (mgb2, {}, dict(extrap=False), 1e-10),
    "slab": (mgb2, {}, dict(selffield_bc="slab"), 1e-10),
    "rkl": (mgb2, {}, dict(rkl=True), 1e-10),
    "pcg_fast_f32": (mgb2, dict(buf=2), dict(sigma_vac=0.01), 1e-7),
    "pcg_tight_f32": (mgb2, dict(buf=2), dict(sigma_vac=0.01, pcg_tol=1e-12), 1e-10),
    "pcg_exact": (mgb2, dict(buf=2), dict(sigma_vac=0.01, pcg_fast=False, pcg_f32=False, pcg_tol=1e-12), 1e-10),
    "explicit_padding": (mgb2, dict(buf=2), dict(sigma_vac=0.01, implicit=False), 1e-10),
    "explicit": (mgb2, {}, dict(implicit=False), 1e-10),
    "three_frustrated": (three_frustrated, {}, {}, 1e-10),
    "kwt10": (lambda: mgb2(gkwt=10.0), {}, {}, 1e-10),
    "kwt10_euler": (lambda: mgb2(gkwt=10.0), {}, dict(rkl=False), 1e-10),
    "kwt10_pcg": (lambda: mgb2(gkwt=10.0), dict(buf=2), dict(sigma_vac=0.01, pcg_tol=1e-12), 1e-10),
    "three_aniso": (three_aniso, {}, {}, 1e-10),
    "hard": (mgb2, {}, dict(hard=True), 1e-10),
    "qp_two": (mgb2, {}, dict(qp_exchange=1.0), 1e-10),
    "qp_two_kwt_pcg": (lambda: mgb2(gkwt=10.0), dict(buf=2), dict(sigma_vac=0.01, pcg_tol=1e-12, qp_exchange=0.5), 1e-10),
    "qp_three": (three_frustrated, {}, dict(qp_exchange=qp_rates(0.5)), 1e-10),
    # x and y extents of the separable solve above 64 and 32: the one-round Float32 x-y launch
    "pcg_wide_f32": (mgb2, dict(nx=68, ny=34, nz=4, buf=2), dict(sigma_vac=0.01), 1e-7),
    # the FP64 x-y launch of the one-round kernel: the separable solve and the exact PCG on wide planes
    "sep_wide": (mgb2, dict(nx=68, ny=34, nz=4), {}, 1e-10),
    # FP64 planes of 141 x 81: the one-round kernel with one m tile a block (two do not fit the LDS)
    "sep_wide160": (mgb2, dict(nx=140, ny=80, nz=4), {}, 1e-10),
    # FP64 planes of 49 x 41: the one-round kernel with 64-row planes
    "sep_mid": (mgb2, dict(nx=48, ny=40, nz=4), {}, 1e-10),
    # 21 planes deep: the FP64 quarter z pass with several modes a wave
    "sep_deep": (mgb2, dict(nx=40, ny=28, nz=20), {}, 1e-10),
    "pcg_exact_wide": (mgb2, dict(nx=68, ny=34, nz=4, buf=2), dict(sigma_vac=0.01, pcg_fast=False, pcg_f32=False,
                                                                  pcg_tol=1e-12), 1e-10),
    # quarters of 71 x 37 (wider than 64 and 32): the folded preconditioner's one-round x-y kernels
    "pcg_mf1_f32": (mgb2, dict(nx=136, ny=68, nz=4, buf=2), dict(sigma_vac=0.01), 1e-7),
    "kwt10_mf1_pcg": (lambda: mgb2(gkwt=10.0), dict(nx=136, ny=68, nz=4, buf=2), dict(sigma_vac=0.01, pcg_tol=1e-12),
                      1e-10),
    "open": (mgb2, {}, dict(selffield_bc="open", open_every=0.02), 1e-10),
    "open_kwt_pcg": (lambda: mgb2(gkwt=10.0), dict(buf=2), dict(sigma_vac=0.01, pcg_tol=1e-12, selffield_bc="open",
                                                                open_every=0.03), 1e-10),
}


def make_case(name, **extra):
    factory, mesh, scheme, tol = CASES[name]
    md = factory()
    m = dict(SMALL)
    m.update(mesh)
    gr = Grid.make(m["nx"], m["ny"], m["nz"], dx=m["dx"], buf=m.get("buf", 1))
    mk = wedge_masks(gr)
    kw = {k: (v(md) if callable(v) else v) for k, v in scheme.items()}
    kw.update(extra)
    st = make_stepper(md, gr, mk, **kw)
    return st, tol


def maxrel(a, b):
    d = np.max(np.abs(a - b)) if a.size else 0.0
    m = np.max(np.abs(b)) if b.size else 0.0
    return d if m == 0 else d / m


def compare_states(sg, s):
    """Per-field errors of compare_cpu.jl:88-96: Ψ_µ against its own size, θ, h and J against
    the size of their whole vector."""
    out = {}
    for mu in range(s.N):
        out[f"G{mu + 1}"] = maxrel(sg.G[mu], s.G[mu])
        out[f"F{mu + 1}"] = maxrel(sg.F[mu], s.F[mu])
    for qs in (("thx", "thy", "thz"), ("hx", "hy", "hz"), ("Jx", "Jy", "Jz"), ("phx", "phy", "phz")):
        m = max(np.max(np.abs(getattr(s, q))) for q in qs)
        for q in qs:
            d = np.max(np.abs(getattr(sg, q) - getattr(s, q)))
            out[q] = d if m == 0 else d / m
    for q in ("Vx", "Vy", "Vz", "Ux", "Uy", "Uz"):
        out[q] = np.max(np.abs(getattr(sg, q) - getattr(s, q)))
    return out


def seed_state(st, H=0.03, J=0.3, warm_steps=5):
    """initial_state and `warm_steps` oracle steps off the seed (compare_cpu.jl:74-75)."""
    from oracle.oracle import Oracle
    s0 = initial_state(st.model, st.grid, st.masks, H)
    HJ = st.selffield(J)
    o = Oracle(st)
    o.set_state(s0)
    if warm_steps:
        o.step(warm_steps, H, HJ)
    s = o.get_state()
    pcg = o.get_pcg() if st.field_mode == 2 else None
    return s, pcg, H, HJ
