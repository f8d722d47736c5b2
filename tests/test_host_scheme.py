"""Host-side Stepper derivation (juliatdgl_amd.scheme) against direct restatements: the bar
self-field tables by windowed matrix sums against the point-by-point midpoint quadrature of
field.jl:151-180."""
import numpy as np
import pytest

from juliatdgl_amd.scheme import Grid, bar_field, bar_field_direct


@pytest.mark.parametrize("dims", [(12, 8, 4, 1), (24, 12, 6, 2), (20, 10, 8, 3)])
def test_bar_field_windowed_sums_are_the_quadrature(dims):
    gr = Grid.make(dims[0], dims[1], dims[2], dx=0.5, buf=dims[3])
    a, b = bar_field(gr), bar_field_direct(gr)
    for k in b:
        assert a[k].shape == b[k].shape
        assert np.max(np.abs(a[k] - b[k])) <= 1e-13 * max(1.0, np.max(np.abs(b[k]))), k
