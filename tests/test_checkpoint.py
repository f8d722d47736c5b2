"""Checkpoints (juliatdgl_amd.checkpoint) on the CPU: the reference's keys, resuming = continuing
(gtdgl/test/test_io.jl:3-16, on the oracle and the golden k25_R4_prod), and the Octave text
checkpoint of the reference's archive (test_io.jl:17-22; data/reference_12x8x4/restart)."""
import os

import numpy as np
import pytest

from juliatdgl_amd.checkpoint import (checkpoint_path, load_state, pcg_extra, read_state, restore_pcg,
                                      save_state)
from juliatdgl_amd.scheme import Grid
from juliatdgl_amd.state import emiphi
from goldens import load_case, same, state_from, stepper_12
from cases import make_case, seed_state

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_save_and_load_keep_the_reference_keys(tmp_path):
    st, _ = make_case("three_frustrated")
    s0, _, _, _ = seed_state(st)
    path = str(tmp_path / "restart" / "state_J+.npz")
    save_state(path, s0, jp=3, J=[0.0, 0.1, 0.2], conv="FFV", confirmed=True, H=0.03)
    assert not os.path.exists(path + ".tmp")
    w = read_state(path)
    assert w["tdgl_version"].startswith("TDGL") and w["nbands"] == 3 and w["jp"] == 3
    assert w["conv"] == "FFV" and w["confirmed"] is True and list(w["J"]) == [0.0, 0.1, 0.2]
    keys = {"G1", "G2", "G3", "thetax", "thetay", "thetaz", "thetax_prev", "thetay_prev", "thetaz_prev",
            "hx", "hy", "hz"}
    assert keys <= set(w)
    s, _ = load_state(path, st.grid)
    for mu in range(3):
        assert np.array_equal(s.G[mu], s0.G[mu])
    for k in ("thx", "thy", "thz", "phx", "phy", "phz", "hx", "hy", "hz"):
        assert np.array_equal(getattr(s, k), getattr(s0, k)), k
    # the links are rebuilt from the phases (state.jl:69-79)
    assert np.array_equal(s.Vx, emiphi(s0.thx)) and np.array_equal(s.Uz, emiphi(s0.phz))
    assert checkpoint_path(str(tmp_path / "restart"), "J+") == path
    with pytest.raises(ValueError, match="the grid needs"):
        load_state(path, Grid.make(10, 8, 4, dx=0.5))


def test_resuming_is_continuing_on_the_golden_case(tmp_path):
    """test_io.jl:3-16: N1 steps in one go, or N1/2, a checkpoint, a load and the rest, give the
    same state; the continuation reaches the golden G2 (1e-10 of max(1, |x|))."""
    from oracle.oracle import Oracle
    par, ini, fin, mk = load_case("k25_R4_prod")
    st = stepper_12(par, mk)
    n1 = int(par["N1"])
    a = Oracle(st)
    a.set_state(state_from(st.grid, ini))
    a.step(n1, par["H"], par["HJ"])
    sa = a.get_state()
    b = Oracle(st)
    b.set_state(state_from(st.grid, ini))
    b.step(n1 // 2, par["H"], par["HJ"])
    path = str(tmp_path / "state_test.npz")
    save_state(path, b.get_state(), jp=3, J=par["J"])
    sc, w = load_state(path, st.grid)
    assert w["jp"] == 3 and w["nbands"] == 2
    c = Oracle(st)
    c.set_state(sc)
    c.step(n1 - n1 // 2, par["H"], par["HJ"])
    sc = c.get_state()
    for x, y in ((sc.G[0], sa.G[0]), (sc.G[1], sa.G[1]), (sc.thx, sa.thx), (sc.hz, sa.hz), (sc.G[1], fin["G2"])):
        ok, d = same(x, y)
        assert ok, d


class _OracleDev:
    """The oracle behind the device stepper's checkpoint calls."""

    def __init__(self, o, st):
        self.o, self.st = o, st

    def get_pcg(self):
        return self.o.get_pcg()

    def set_pcg(self, p):
        self.o.set_pcg(p)


def test_pcg_history_travels_with_the_checkpoint(tmp_path):
    """The warm PCG solution and the one before (protocol.jl:656-658, 380-381): a resumed PCG run
    starts from the same guess, so it takes the same iterations to the same state."""
    from oracle.oracle import Oracle
    st, _ = make_case("pcg_tight_f32")
    s0, pcg, H, HJ = seed_state(st)
    a = Oracle(st)
    a.set_state(s0)
    a.set_pcg(pcg)
    a.step(6, H, HJ)
    ex = pcg_extra(_OracleDev(a, st))
    assert {"pcg_x1", "pcg_x3", "pcg_p2"} <= set(ex)
    path = save_state(str(tmp_path / "state_pcg.npz"), a.get_state(), **ex)
    a.step(4, H, HJ)
    ref = a.get_state()
    s, w = load_state(path, st.grid)
    b = Oracle(st)
    b.set_state(s)
    restore_pcg(_OracleDev(b, st), w)
    b.step(4, H, HJ)
    got = b.get_state()
    for mu in range(st.N):
        assert np.max(np.abs(got.G[mu] - ref.G[mu])) <= 1e-12
    assert np.max(np.abs(got.hz - ref.hz)) <= 1e-12
    assert b.get_pcg().iters == a.get_pcg().iters


def test_octave_text_checkpoint_of_the_reference_archive():
    """test_io.jl:17-22: a release-1.1 Octave text checkpoint (data/reference_12x8x4/restart/
    state_Jpos.mat) loads on the 12x8x4 grid; without previous phases θ_prev = θ."""
    path = os.path.join(GOLDEN, "octave_state_Jpos.mat")
    raw = read_state(path)
    assert raw["jp"] == 1.0 and raw["G1"].shape == (15, 11, 7) and raw["thetax"].shape == (14, 11, 7)
    gr = Grid.make(12, 8, 4, dx=0.5)
    s, w = load_state(path, gr)
    assert np.array_equal(s.G[0], raw["G1"]) and np.array_equal(s.thz, raw["thetaz"])
    assert np.array_equal(s.phx, s.thx)
    # the first interior value of the file, as written: (0.74409940456024903, 0.59597626746806065)
    nz = s.G[0][np.abs(s.G[0]) > 0]
    assert nz[0] == complex(0.74409940456024903, 0.59597626746806065)
    assert np.all(np.isfinite(s.G[1])) and np.max(np.abs(s.G[0])) < 1.5
