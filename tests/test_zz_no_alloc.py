"""tdgl_step_noalloc (the entry point that allocates nothing, its reductions' scratch in the
library's fixed device array) against tdgl_step: the same bits, on two calls in a row, the second
on a smaller grid (the array keeps the first call's contents).  Needs a HIP device."""
import numpy as np
import pytest

from cases import make_case, mgb2, seed_state

from juliatdgl_amd.scheme import Grid, make_stepper, wedge_masks

pytestmark = pytest.mark.gpu


def _run(st, s0, pcg, H, HJ, noalloc, n):
    from juliatdgl_amd.hip import HipStepper
    dev = HipStepper(st)
    dev.upload(s0)
    if pcg is not None:
        dev.set_pcg(pcg)
    (dev.step_noalloc if noalloc else dev.step)(n, H, HJ)
    out = dev.download()
    p = dev.get_pcg() if pcg is not None else None
    dev.close()
    return out, p


def test_step_noalloc_equals_step():
    from juliatdgl_amd.hip import device_count
    if device_count() < 1:
        pytest.fail("no HIP device visible")
    big, _ = make_case("kwt10_pcg")  # 24x12x6 cells, buf 2: the PCG with the Float32 preconditioner
    gr = Grid.make(12, 8, 4, dx=0.5, buf=2)
    small = make_stepper(mgb2(), gr, wedge_masks(gr), sigma_vac=0.01, pcg_fast=False, pcg_f32=False, pcg_tol=1e-12)
    cases = [big, small]
    assert small.field_mode == 2 and small.grid.nodes < big.grid.nodes
    for st in cases:
        s0, pcg, H, HJ = seed_state(st)
        print(f"[gpu] tdgl_step_noalloc on {st.grid.Nx}x{st.grid.Ny}x{st.grid.Nz}", flush=True)
        ref, pref = _run(st, s0, pcg, H, HJ, False, 12)
        got, pgot = _run(st, s0, pcg, H, HJ, True, 12)
        for (k, a), (_, b) in zip(ref.arrays(), got.arrays()):
            assert np.array_equal(a, b), k
        assert pgot.iters == pref.iters and pgot.warm == pref.warm and pgot.have2 == pref.have2
        for a, b in zip(pref.x, pgot.x):
            assert np.array_equal(a, b)
