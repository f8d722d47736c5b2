"""A whole two-polarity ladder against release 1.2's golden one (gtdgl/test/test_protocol.jl:4-29,
golden ladder_k25.jld2): the κ = 25 notched box, currents 0.3 .. 1.5 and the seed point, point by
point — the verdicts and cycles exactly, the brackets exactly, the voltages, slips, magnetization and
residuals and the final Ψ to 1e-10, on the oracle here (the device: tests/test_gpu_meter.py)."""
import json
import os

import numpy as np
import pytest

from juliatdgl_amd.meter import Protocol, eps_cycle, run_ladders
from juliatdgl_amd.scheme import Band, Grid, Model, make_stepper, wedge_masks

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

# the golden fields, each to 1e-10 of max(1, |x|), the reference's bar between its two CPU releases
# (test_protocol.jl:18-24); the verdicts, cycles and brackets exactly
FIELDS = {"V": "V", "V2": "Vb2", "Vst": "Vst", "Vh": "Vh1", "Vh2": "Vh2", "nslip": "nslip", "M": "M", "r1": "r1",
          "r2": "r2"}
TOL = 1e-10


def golden_setup():
    with open(os.path.join(GOLDEN, "ladder_k25_scalars.json")) as fh:
        sc = json.load(fh)
    z = np.load(os.path.join(GOLDEN, "ladder_k25.npz"))
    md = Model.make((Band.make(), Band.make(m=sc["mr"], Gamma=1.0, a=0.59)), [[0.0, 1e-3], [1e-3, 0.0]], kappa=25.0)
    gr = Grid.make(24, 12, 4, dx=0.5)
    st = make_stepper(md, gr, wedge_masks(gr), selffield_bc="slab", rkl=False)
    pr = Protocol(t_min_GL=20.0, t_min_GL_res=60.0, mMax=4000, delta_v=6, v_margin=1.0, f_strict=False)
    return sc, z, st, pr


def _vals(L, key):
    return {"V": L.V, "Vb2": L.Vb[:, 1], "Vst": L.Vst, "Vh1": L.Vh[:, 0], "Vh2": L.Vh[:, 1], "nslip": L.nslip,
            "M": L.M, "r1": L.res[:, 0], "r2": L.res[:, 1]}[key]


def check_golden_ladder(backend_of, report=print):
    sc, z, st, pr = golden_setup()
    assert st.dt == sc["dt"] and st.nsub == (1, sc["n_pi"])
    assert eps_cycle(pr, st) == sc["eps_c"]
    pos, neg, br = run_ladders(backend_of, st, [0.30, 0.60, 0.90, 1.20, 1.50], H=0.03, pr=pr)
    worst = {}
    for L, pol in ((pos, "pos"), (neg, "neg")):
        assert L.J == list(z[f"{pol}_J"])
        assert "".join(L.conv) == bytes(z[f"{pol}_conv"]).decode(), (pol, L.conv)
        assert list(L.m) == list(z[f"{pol}_m"]), (pol, L.m)
        for g, key in FIELDS.items():
            a, b = np.asarray(_vals(L, key), dtype=float), z[f"{pol}_{g}"]
            d = float(np.max(np.abs(a - b) / np.maximum(1.0, np.abs(b))))
            worst[f"{pol}_{g}"] = d
            assert d <= TOL, (pol, g, d, a, b)
        for mu, g in enumerate(("G1", "G2")):
            b = z[f"{pol}_{g}"]
            d = float(np.max(np.abs(L.state.G[mu] - b) / np.maximum(1.0, np.abs(b))))
            worst[f"{pol}_{g}"] = d
            assert d <= TOL, (pol, g, d)
    report("    worst " + ", ".join(f"{k} {v:.1e}" for k, v in sorted(worst.items(), key=lambda kv: -kv[1])[:6]))
    assert (br["pos"]["lo"], br["pos"]["hi"], br["neg"]["lo"], br["neg"]["hi"]) == \
        (sc["lo_pos"], sc["hi_pos"], sc["lo_neg"], sc["hi_neg"])
    return worst


def test_golden_ladder_on_the_oracle():
    from oracle.meter import OracleBackend
    from oracle.oracle import Oracle

    def backend_of(st):
        return _OracleLadderBackend(OracleBackend(Oracle(st), st))

    check_golden_ladder(backend_of)


class _OracleLadderBackend:
    """OracleBackend with the upload / download a ladder uses."""

    def __init__(self, ob):
        self.ob = ob

    def upload(self, s):
        self.ob.o.set_state(s)

    def download(self):
        return self.ob.o.get_state()

    def step(self, n, H, HJ):
        self.ob.step(n, H, HJ)

    def meter_begin(self, vm):
        self.ob.meter_begin(vm)

    def meter_cycle(self, first):
        return self.ob.meter_cycle(first)
