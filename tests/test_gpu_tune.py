"""tdgl_autotune: the state, the PCG history and the open boundary come back as they were after the
tuner's timed steps, every candidate it names was timed, and the steps after a tune agree with the
steps of an untuned handle.  The PCG box kernels' levels per block decide how the reductions' block
partials are summed and the preconditioner's folds how it rounds, so the two agree to the case's
tolerance against the oracle, not bit for bit; the m tiles, the z pass's waves and its columns or
x-runs change no bit (test_gpu_parity's kernel-variant test).  Needs a HIP device."""
import numpy as np
import pytest

from cases import compare_states, make_case, seed_state

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["kwt10_pcg", "pcg_wide_f32", "open_kwt_pcg", "qp_two_kwt_pcg"])
def test_tune_restores_state_and_steps_agree(name):
    from juliatdgl_amd.hip import HipStepper, device_count
    if device_count() < 1:
        pytest.fail("no HIP device visible")
    st, tol = make_case(name)
    s0, pcg, H, HJ = seed_state(st)
    print(f"[gpu] tdgl_autotune on {name}", flush=True)
    a = HipStepper(st)
    b = HipStepper(st)
    try:
        for d in (a, b):
            d.upload(s0)
            d.set_pcg(pcg)
        a.step(2, H, HJ)  # F and the normal fluids' potentials of a step
        b.step(2, H, HJ)
        before, pb = a.download(), a.get_pcg()
        qb = a.get_qp() if st.qp is not None else []
        t = a.autotune(H, HJ)
        after, pa = a.download(), a.get_pcg()
        for x, y in zip(qb, a.get_qp() if st.qp is not None else []):
            assert np.array_equal(x, y)
        for (k, x), (_, y) in zip(before.arrays(), after.arrays()):
            assert np.array_equal(x, y), k
        assert (pa.iters, pa.warm, pa.have2) == (pb.iters, pb.warm, pb.have2)
        for x, y in zip(pb.x + pb.xprev, pa.x + pa.xprev):
            assert np.array_equal(x, y)
        timed = {}
        for kind, value, us in t["candidates"]:
            assert us > 0
            timed.setdefault(kind, set()).add(value)
        assert t["box_kc"] in timed["box_kc"] and t["plane_mt"] in timed["plane_mt"]
        if "zmodes_waves" in timed:
            assert t["zmodes_waves"] in timed["zmodes_waves"]
        if "qparts" in timed:
            assert t["qparts"] in timed["qparts"] and timed["qparts"] == {1, 2}
        else:
            assert t["qparts"] == 1
        # the preconditioner's path: the folds (timed where the handle has them) or the full-length
        # transforms with one of the two z passes
        if t["fold"]:
            assert 1 in timed["fold"]
        else:
            assert t["zcol"] in timed["zcol"]
        assert a.stats()["folded"] == t["fold"]
        a.step(12, H, HJ)
        b.step(12, H, HJ)
        errs = compare_states(a.download(), b.download())
        worst = max(errs.values())
        print(f"[gpu]   tuned {t['box_kc'], t['plane_mt'], t['zmodes_waves'], t['fold'], t['zcol']}: worst {worst:.2e}",
              flush=True)
        assert worst <= tol, errs
    finally:
        a.close()
        b.close()
