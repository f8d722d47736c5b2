"""Host-side Stepper derivation (juliatdgl_amd.scheme) against direct restatements: the bar
self-field tables by windowed matrix sums against the point-by-point midpoint quadrature of
field.jl:151-180, and that quadrature against the closed-form integral it approximates."""
import math

import numpy as np
import pytest

from juliatdgl_amd.scheme import Grid, bar_field, bar_field_direct


def _bar_closed_form(gr):
    """The 2D Biot–Savart field of a uniform current in the sample's y-z cross-section (the
    integral field.jl:151-180 approximates), per unit HJ, in closed form: with u = y − Y and
    v = z − Z, ∫∫ u/(u² + v²) dY dZ is the corner sum of G(u, v) = (v/2) ln(u² + v²) + u·atan(v/u) − v
    (∂²G/∂u∂v = u/(u² + v²)) and ∫∫ v/(u² + v²) that of G(v, u); hz = c ∫∫ u/r², hy = −c ∫∫ v/r² with
    c = −2/d / 2π."""
    def G(u, v):
        r2 = u * u + v * v
        with np.errstate(divide="ignore", invalid="ignore"):
            t1 = np.where(v == 0, 0.0, 0.5 * v * np.log(np.where(r2 > 0, r2, 1.0)))
            t2 = np.where(u == 0, 0.0, u * np.arctan(v / np.where(u == 0, 1.0, u)))
        return t1 + t2 - v

    ya, yb, za, zb = (gr.y1 - 1) * gr.dy, gr.y2 * gr.dy, (gr.z1 - 1) * gr.dz, gr.z2 * gr.dz
    c = (-2 / ((gr.z2 - gr.z1 + 1) * gr.dz)) / (2 * math.pi)

    def corners(F, y, z):
        u1, u2, v1, v2 = y - yb, y - ya, z - zb, z - za
        return F(u2, v2) - F(u1, v2) - F(u2, v1) + F(u1, v1)

    Ny, Nz = gr.Ny, gr.Nz
    # hy at (y nodes, z centres), hz at (y centres, z nodes) of the faces
    Y, Z = np.meshgrid(np.arange(Ny + 1) * gr.dy, (np.arange(Nz) + 0.5) * gr.dz, indexing="ij")
    hy = -c * corners(lambda u, v: G(v, u), Y, Z)
    Y, Z = np.meshgrid((np.arange(Ny) + 0.5) * gr.dy, np.arange(Nz + 1) * gr.dz, indexing="ij")
    hz = c * corners(G, Y, Z)
    return hy, hz


@pytest.mark.parametrize("dims", [(12, 8, 4, 1), (24, 12, 6, 2), (16, 10, 4, 1)])
def test_bar_field_is_the_closed_form_integral(dims):
    """The bar profiles (the midpoint quadrature of field.jl:151-180, nq = 8 points per cell) against
    the exact integral of the uniform current over the cross-section: within 1e-8 of the field's
    maximum, and 10 times closer at nq = 16 (a sign, a factor or a swapped axis would be O(1))."""
    gr = Grid.make(*dims[:3], dx=0.5, buf=dims[3])
    hy, hz = _bar_closed_form(gr)
    errs = []
    for nq in (8, 16):
        b = bar_field(gr, nq)
        errs.append(max(np.max(np.abs(b["hy_x"] - hy)) / np.max(np.abs(hy)),
                        np.max(np.abs(b["hz_x"] - hz)) / np.max(np.abs(hz))))
    assert errs[0] < 1e-8, errs
    assert errs[1] < errs[0] / 10, errs


@pytest.mark.parametrize("dims", [(12, 8, 4, 1), (24, 12, 6, 2), (20, 10, 8, 3)])
def test_bar_field_windowed_sums_are_the_quadrature(dims):
    gr = Grid.make(dims[0], dims[1], dims[2], dx=0.5, buf=dims[3])
    a, b = bar_field(gr), bar_field_direct(gr)
    for k in b:
        assert a[k].shape == b[k].shape
        assert np.max(np.abs(a[k] - b[k])) <= 1e-13 * max(1.0, np.max(np.abs(b[k]))), k
