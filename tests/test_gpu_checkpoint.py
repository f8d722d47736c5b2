"""Checkpoints and snapshots of the device stepper (juliatdgl_amd.checkpoint): a run stopped at a
checkpoint and resumed on a fresh stepper continues it (gpu/TDGLGPU/test/test_resume.jl:29-52: to
1e-7 of the larger field, its GPU bound), and a device snapshot continues it to the bit.  Needs a
HIP device."""
import numpy as np
import pytest

from juliatdgl_amd.checkpoint import (load_snapshot, restore, resume_checkpoint, save_checkpoint,
                                      save_snapshot, snapshot)
from cases import make_case, seed_state

pytestmark = pytest.mark.gpu

FIELDS = ("thx", "thy", "thz", "phx", "phy", "phz", "hx", "hy", "hz")


def _hip():
    from juliatdgl_amd.hip import HipStepper, device_count
    if device_count() < 1:
        pytest.fail("no HIP device visible")
    return HipStepper


def _state_error(a, b):
    """test_resume.jl:44-46: the largest difference over the bands and fields, each against the
    larger of the two fields (and 1)."""
    arrs = [(x, y) for x, y in zip(a.G, b.G)] + [(getattr(a, f), getattr(b, f)) for f in FIELDS]
    return max(np.max(np.abs(x - y)) / max(np.max(np.abs(x)), np.max(np.abs(y)), 1.0) for x, y in arrs)


@pytest.mark.parametrize("name", ["production", "slab", "kwt10_pcg", "open_kwt_pcg"])
def test_device_resume_from_a_checkpoint_continues_the_run(name, tmp_path):
    HipStepper = _hip()
    st, _ = make_case(name)
    s0, pcg, H, HJ = seed_state(st)
    print(f"[gpu] checkpoint resume, {name}", flush=True)
    a = HipStepper(st)
    a.upload(s0)
    if pcg is not None:
        a.set_pcg(pcg)
    a.step(10, H, HJ)
    path = save_checkpoint(str(tmp_path / "restart" / "state_J+.npz"), a, jp=1, J=[0.1])
    a.step(10, H, HJ)
    sa = a.download()
    b = HipStepper(st)
    w = resume_checkpoint(path, b)
    assert w["jp"] == 1
    b.step(10, H, HJ)
    err = _state_error(sa, b.download())
    print(f"    state error {err:.3e}", flush=True)
    assert err <= 1e-7
    if st.field_mode == 2:
        xa, xb = a.get_pcg().x, b.get_pcg().x
        assert max(np.max(np.abs(p - q)) for p, q in zip(xa, xb)) <= 1e-7
    if st.selffield_bc == "open":
        ca, oa, na = a.get_open()
        cb, ob, nb = b.get_open()
        assert ca == cb
        assert all(np.allclose(p, q, rtol=1e-7, atol=1e-7) for p, q in zip(oa + na, ob + nb))


@pytest.mark.parametrize("name", ["kwt10_pcg", "open_kwt_pcg"])
def test_device_snapshot_continues_to_the_bit(name, tmp_path):
    HipStepper = _hip()
    st, _ = make_case(name)
    s0, pcg, H, HJ = seed_state(st)
    print(f"[gpu] snapshot, {name}", flush=True)
    a = HipStepper(st)
    a.upload(s0)
    if pcg is not None:
        a.set_pcg(pcg)
    a.step(7, H, HJ)
    path = save_snapshot(str(tmp_path / "snap.npz"), snapshot(a))
    a.step(9, H, HJ)
    sa = a.download()
    b = HipStepper(st)
    restore(b, load_snapshot(path, b))
    b.step(9, H, HJ)
    sb = b.download()
    for x, y in zip(sa.G + sa.F, sb.G + sb.F):
        assert np.array_equal(x, y)
    for f in FIELDS + ("Vx", "Vy", "Vz", "Ux", "Uy", "Uz", "Jx", "Jy", "Jz"):
        assert np.array_equal(getattr(sa, f), getattr(sb, f)), f
    assert a.get_pcg().iters == b.get_pcg().iters
