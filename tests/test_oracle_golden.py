"""The oracle against the reference's golden states: gtdgl/test/test_regression12.jl:59-80.

juliaTDGL 1.2 wrote "ini" after 200 steps and "fin" after 40 more; 40 steps of the release-2.0
step from "ini" must reproduce "fin" to 1e-10 of max(1, |x|) in every field.  The three cases
pin the Euler band pass, the π-band Euler substeps with extrapolated link roots (_rot!,
_advance_links!), the supercurrent, the field increment, the separable implicit solve (the two
"prod" cases), the explicit curl-curl ("explicit"), links, curl, the slab field boundary, the
Neumann ghosts and the advance; and the Stepper's dt and substep count.
"""
import numpy as np
import pytest

from goldens import CASES, load_case, same, state_from, stepper_12
from oracle.oracle import Oracle


@pytest.mark.parametrize("name", CASES)
def test_stepper_matches_golden_parameters(name):
    par, ini, fin, mk = load_case(name)
    st = stepper_12(par, mk)
    assert st.dt == par["dt"]
    assert st.nsub == (1, par["n_pi"])
    assert st.selffield(par["J"]) == par["HJ"]


@pytest.mark.parametrize("name", CASES)
def test_oracle_replays_golden(name):
    par, ini, fin, mk = load_case(name)
    st = stepper_12(par, mk)
    s = state_from(st.grid, ini)
    for k in ("Vx", "Ux", "Uz"):
        ok, d = same(getattr(s, k), ini[k])
        assert ok, (k, d)
    o = Oracle(st)
    o.set_state(s)
    o.step(par["N1"], par["H"], par["HJ"])
    out = o.get_state()
    for k, ref in (("G1", fin["G1"]), ("G2", fin["G2"])):
        ok, d = same(out.G[int(k[1]) - 1], ref)
        assert ok, (k, d)
    for k in ("thetax", "thetay", "thetaz", "phix", "phiy", "phiz"):
        mine = getattr(out, {"theta": "th", "phi": "ph"}[k[:-1]] + k[-1])
        ok, d = same(mine, fin[k])
        assert ok, (k, d)
    for k in ("Vx", "Vy", "Vz", "hx", "hy", "hz"):
        ok, d = same(getattr(out, k), fin[k])
        assert ok, (k, d)


def test_golden_changes_state():
    """A planted mistake must fail: 40 steps move the state by far more than the tolerance."""
    par, ini, fin, mk = load_case("k2_R4_prod")
    ok, d = same(ini["G2"], fin["G2"])
    assert not ok and d > 1e-6
