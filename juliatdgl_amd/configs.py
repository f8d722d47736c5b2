"""The benchmark workloads: the stepper configurations of the measured runs, built from the
material and mesh descriptions they are quoted on.

Conventions (the reference's, gtdgl/src/model.jl and config.jl:32-50): condensation coefficient
a > 0 for a condensed band (a band quoted as a = -1 in the α convention is a = 1 here); a band
with coherence length ratio ξ_µ/ξ_1 = R has mass m = 1/(R²·a/β); u = Γ/σn.
"""
from __future__ import annotations

import math

from .scheme import Band, Grid, Model, make_stepper, notch_masks, sloped_wedge_masks, two_band, wedge_masks


def two_band_model(kappa=2.5, Gamma=1.0, gkwt=0.0):
    """Two bands: (a, β, ξ, Γ) = (1, 1, 1, Γ) and (0.5, 1, 2.5, Γ), Josephson −0.1, κ = 2.5."""
    a2, R2 = 0.5, 2.5
    return Model.make((Band.make(m=1.0, Gamma=Gamma, a=1.0, gkwt=gkwt),
                       Band.make(m=1 / (R2 * R2 * (a2 / 1.0)), Gamma=Gamma, a=a2, gkwt=gkwt)),
                      [[0.0, -0.1], [-0.1, 0.0]], kappa=kappa, sigma_n=1.0)


def notch_12x8x4():
    """12x8x4 cells, buf 1, dx 0.5; notch f = 0 on cells x in [5,7), y in [5,8); bar field,
    KWT and noise off, dt = 0.01, separable implicit field.  H = 0.3, J = 0.1."""
    md = two_band_model()
    gr = Grid.make(12, 8, 4, dx=0.5, buf=1)
    mk = notch_masks(gr, (5, 7), (5, 8))
    st = make_stepper(md, gr, mk, dt=0.01, selffield_bc="bar")
    return st, dict(H=0.3, J=0.1)


U_KWT = 5.79       # u = Γ/σn of a dirty BCS band (model.jl:52)
GKWT = 10.0        # KWT γ̃ (the reference's example value, gtdgl/examples/kwt_k2.toml)


def kwt_pcg_128x64x16(nx=128, ny=64, nz=16, pcg_tol=1e-10, J=0.1, dt_scale=0.1):
    """The two bands of notch_12x8x4 (Γ = 1) with KWT: u = Γ/σn = 5.79 (σn = 1/5.79) and γ̃ = 10, on
    128x64x16 cells, dx 0.5, an insulating buffer of 2 cells (σvac/σn = 0.01: the PCG solve) to
    relative residual 1e-10, a wedge cut into face y1 whose sloped face is at 30 degrees, bar
    field, H = 0.2.  The stepper's own dt rule with dt_scale = 0.1: at the default 0.5 (and at 0.25)
    the staged KWT bands diverge within about a dozen steps, with either reading of u (Γ = 1 and
    σn = 1/5.79, or Γ = 5.79 and σn = 1), and it is the rotating frame that does (the same step
    without the frame, or without KWT, stays bounded); at 0.1 the run is stable.  J is one rung of
    the 0 → 0.3 ramp."""
    sn = 1 / U_KWT
    base = two_band_model(Gamma=1.0, gkwt=GKWT)
    md = Model.make(base.bands, [[0.0, -0.1], [-0.1, 0.0]], kappa=2.5, sigma_n=sn)
    gr = Grid.make(nx, ny, nz, dx=0.5, buf=2)
    mk = sloped_wedge_masks(gr, angle_deg=30.0, depth_frac=0.35, x_frac=0.45)
    st = make_stepper(md, gr, mk, sigma_vac=0.01 * sn, pcg_tol=pcg_tol, selffield_bc="bar", dt_scale=dt_scale)
    return st, dict(H=0.2, J=J)


def three_band_aniso_model(kappa=2.5):
    """Three bands (a = 1, 0.6, 0.3; ξ = 1, 1.8, 3; Josephson −0.1, −0.05, −0.02) with
    m_z = 4 m_xy (m_axes = (m, m, 4m), m = 1/(ξ² a))."""
    bands = []
    for a, xi in ((1.0, 1.0), (0.6, 1.8), (0.3, 3.0)):
        m = 1 / (xi * xi * a)
        bands.append(Band.make(a=a, m_axes=(m, m, 4 * m)))
    return Model.make(tuple(bands), [[0, -0.1, -0.05], [-0.1, 0, -0.02], [-0.05, -0.02, 0]], kappa=kappa)


def open_3band_256x128x32(nx=256, ny=128, nz=32, buf=8, pcg_tol=1e-8, selffield_bc="open", **kw):
    """The three anisotropic bands of three_band_aniso_model on 256x128x32 cells and an insulating
    buffer of 8 cells (σvac = 0.01, PCG with the Float32 preconditioner and the FP64 outer
    iteration), the open boundary (h_δ by the zero-padded FFT convolution, recomputed every t_GL),
    a notch at the middle of face y1."""
    md = three_band_aniso_model()
    gr = Grid.make(nx, ny, nz, dx=0.5, buf=buf)
    mk = notch_masks(gr, (nx // 2 - 4, nx // 2 + 4), (0, ny // 4))
    st = make_stepper(md, gr, mk, sigma_vac=0.01, pcg_tol=pcg_tol, selffield_bc=selffield_bc, **kw)
    return st, dict(H=0.2, J=0.1)


def pcg_wedge_96x48x24():
    """The reference's validation run file gpu/TDGLGPU/validation/configs/pcg96.toml: material
    mgb2_model_II (σ band and a π band of ξ ratio 4, a = 0.59, γ12 = 1e-3) at κ = 2, 96x48x24
    cells, dx 0.5, buffer 2 with σvac = 0.01 (the insulating PCG with the scheme defaults: tol
    1e-8, extrapolated first guess, Float32 preconditioner), the catalogue WEDGE, J = 0.3, H = 0.03."""
    md = two_band(xi_ratio=4.0, kappa=2.0)
    gr = Grid.make(96, 48, 24, dx=0.5, buf=2)
    st = make_stepper(md, gr, wedge_masks(gr), sigma_vac=0.01)
    return st, dict(H=0.03, J=0.3)


def plain_wedge_96x48x24():
    """validation/configs/plain96.toml: as pcg_wedge_96x48x24 with a one-cell normal-metal buffer
    (the separable implicit solve)."""
    md = two_band(xi_ratio=4.0, kappa=2.0)
    gr = Grid.make(96, 48, 24, dx=0.5, buf=1)
    st = make_stepper(md, gr, wedge_masks(gr))
    return st, dict(H=0.03, J=0.3)
