"""A run file driven end to end on the device (gtdgl/src/config.jl:174-242): the TOML file with its
material, mesh, a shape file next to it (CUSTOM, and CUSTOM_REF its mirror), scheme, protocol and
run tables, loaded by load_config, stepped by the HIP stepper from its start state, and each of
its ladder points closed with the same verdict after the same number of cycles as the oracle
closes it on the same run file.  Needs a HIP device."""
import numpy as np
import pytest

from juliatdgl_amd.meter import Voltmeter, run_rung
from juliatdgl_amd.runfile import load_config, start_state, stepper
from juliatdgl_amd.state import initial_state

pytestmark = pytest.mark.gpu

SHAPE = '''# a triangle on face y1 and a buried disc
[[shape]]
type = "triangle"
x = 6.0
y = 0.0
w = 4.0
h = 2.5
skew = 1.0
f = 0.0
[[shape]]
type = "circle"
x = 3.0
y = 4.0
r = 0.8
z = "buried"
f = 0.3
'''

RUN = '''[material]
kappa = 25.0
gamma12 = 0.001
[[material.band]]
name = "sigma"
m = 1.0
[[material.band]]
name = "pi"
xi_ratio = 4.0
a = 0.59
[mesh]
nx = 24
ny = 12
nz = 4
dx = 0.5
[geometry]
GEOM_MODE = "{mode}"
geom_file = "shape.toml"
[scheme]
rkl = "rkc"
[protocol]
t_min_GL = 20.0
t_min_GL_res = 60.0
mMax = 400
delta_v = 6
[run]
J = [0.5, 1.4]
H = 0.03
'''


@pytest.mark.parametrize("mode", ["CUSTOM", "CUSTOM_REF"])
def test_run_file_on_the_device(tmp_path, mode):
    from juliatdgl_amd.hip import HipStepper, device_count
    from oracle.meter import OracleBackend
    from oracle.oracle import Oracle
    if device_count() < 1:
        pytest.fail("no HIP device visible")
    (tmp_path / "shape.toml").write_text(SHAPE)
    (tmp_path / "run.toml").write_text(RUN.format(mode=mode))
    c = load_config(str(tmp_path / "run.toml"))
    assert (c.masks.f < 1).sum() > 10
    st = stepper(c)
    s0 = start_state(c)
    assert s0 is None  # "meissner": the Meissner seed
    s0 = initial_state(c.model, c.grid, c.masks, c.H)
    pr = c.protocol
    vm = Voltmeter.of_ladder(st, pr.delta_v)
    for J in c.J:
        print(f"[gpu] run file {mode}: rung at J = {J}", flush=True)
        o = Oracle(st)
        o.set_state(s0)
        ob = OracleBackend(o, st)
        dev = HipStepper(st)
        dev.upload(s0)
        out = []
        for b in (ob, dev):
            b.meter_begin(vm)
            out.append(run_rung(b, st, pr, J, c.H))
        dev.close()
        ro, rd = out
        print(f"    oracle {ro.conv} m={ro.m}, device {rd.conv} m={rd.m}", flush=True)
        assert rd.conv == ro.conv and rd.m == ro.m
        assert np.allclose(rd.V, ro.V, rtol=1e-8, atol=1e-8)
