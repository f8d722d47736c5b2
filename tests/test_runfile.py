"""Run files (juliatdgl_amd.runfile) against the reference's: its validation run files give the
steppers built by hand for them, gtdgl/test/test_config.jl's run file gives the golden ladder's
parameters (ladder_k25.jld2: mr, dt, n_pi), the catalogue masks are the ones release 1.2 wrote
(gtdgl/test/golden/masks.jld2), and the keys of config.jl map as there."""
import json
import os

import numpy as np
import pytest

from juliatdgl_amd import configs
from juliatdgl_amd.checkpoint import save_state
from juliatdgl_amd.runfile import geometry_masks, load_config, material, start_state, stepper
from juliatdgl_amd.scheme import Grid, catalogue_masks
from juliatdgl_amd.state import initial_state

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

TEST_CONFIG_RUN = """
[material]
file = "mgb2_model_II.toml"
[mesh]
nx = 24
ny = 12
nz = 4
dx = 0.5
[geometry]
GEOM_MODE = "WEDGE"
inj_margin = 4
[scheme]
threads = false
selffield_bc = "slab"
rkl = false
[protocol]
t_min_GL = 20.0
t_min_GL_res = 60.0
mMax = 4000
delta_v = 6
v_margin = 1.0
f_strict = false
[run]
J = [0.30, 0.60, 0.90, 1.20, 1.50]
H = 0.03
"""


def _same_stepper(a, b):
    assert a.dt == b.dt and a.rkl == b.rkl and a.nsub == b.nsub and a.field_mode == b.field_mode
    assert a.bc == b.bc and a.bcs == b.bcs and a.divisor == b.divisor and a.selffield_bc == b.selffield_bc
    assert (a.pcg_tol, a.pcg_fast, a.pcg_f32) == (b.pcg_tol, b.pcg_fast, b.pcg_f32)
    for x, y in zip(a.masks.__dict__.values(), b.masks.__dict__.values()):
        assert np.array_equal(x, y)
    assert (a.D is None) == (b.D is None)
    if a.D is not None:
        assert all(np.array_equal(x, y) for x, y in zip(a.D, b.D))


@pytest.mark.parametrize("name,factory", [("pcg96", configs.pcg_wedge_96x48x24), ("plain96", configs.plain_wedge_96x48x24)])
def test_validation_run_files_give_their_steppers(name, factory):
    """gpu/TDGLGPU/validation/configs/{pcg96,plain96}.toml: the Stepper from the file is the one
    built by hand for the benchmark workload."""
    c = load_config(os.path.join(GOLDEN, "runfiles", name + ".toml"))
    st_ref, p = factory()
    _same_stepper(stepper(c), st_ref)
    assert c.J == [p["J"]] and c.H == p["H"]


def test_the_reference_test_run_file(tmp_path):
    """test_config.jl:3-33: the run file's model is the golden ladder's (mr to the bit), and its
    Stepper takes the golden's dt and substeps of the π band."""
    f = tmp_path / "run.toml"
    f.write_text(TEST_CONFIG_RUN)
    c = load_config(str(f))
    with open(os.path.join(GOLDEN, "ladder_k25_scalars.json")) as fh:
        d = json.load(fh)
    assert c.model.bands[1].m == d["mr"]
    st = stepper(c)
    assert st.dt == d["dt"] and st.nsub[1] == d["n_pi"] and st.selffield_bc == "slab" and st.rkl == (0, 0)
    assert c.protocol.t_min_GL == 20.0 and c.protocol.mMax == 4000 and c.protocol.delta_v == 6 and not c.protocol.f_strict
    assert c.J == [0.3, 0.6, 0.9, 1.2, 1.5] and c.H == 0.03 and c.model.kappa == 25.0


def test_catalogue_masks_are_release_12s():
    """The [geometry] table's masks (geometry_masks, the run file's path into juliatdgl_amd.geometry)
    equal the golden ones of every catalogue case to the bit (the shape files: tests/test_geometry.py)."""
    z = np.load(os.path.join(GOLDEN, "catalogue_masks.npz"))
    meta = json.loads(bytes(z["meta_json"]).decode())["cases"]
    n = 0
    for k, m in meta.items():
        ma, g = m["malla"], m["geo"]
        if "geom_file" in g:
            continue
        gr = Grid.make(ma["Nx_sc"], ma["Ny_sc"], ma["Nz_sc"], dx=ma["dx"], buf=ma["buf"])
        mk = geometry_masks(dict(g), gr)
        assert np.array_equal(mk.f, z[k]), k
        n += 1
    assert n == 82
    mk = catalogue_masks(Grid.make(40, 24, 6, dx=0.5, buf=1), "WEDGE_BOTH_REF", wedge_depth2=0.0)
    assert np.array_equal(mk.f, z["WEDGE_BOTH_REF_40x24x6"])
    with pytest.raises(NotImplementedError, match="current_axis"):
        geometry_masks({"GEOM_MODE": "WEDGE", "current_axis": "y"}, Grid.make(8, 6, 4, dx=0.5))


def test_run_file_with_a_shape_file(tmp_path):
    """A run file whose [geometry] names a shape file next to it (config.jl:182-186): its masks are
    the shape's, mirrored for CUSTOM_REF; a missing file is named."""
    (tmp_path / "tri.toml").write_text('[[shape]]\ntype = "triangle"\nx = 8.0\ny = 0.0\nw = 5.0\nh = 3.0\nf = 0.0\n')
    base = ('[material]\nkappa = 2.0\n[[material.band]]\nm = 1.0\n[mesh]\nnx = 24\nny = 12\nnz = 4\ndx = 0.5\n'
            '[run]\nJ = [0.5]\nH = 0.03\n')
    from juliatdgl_amd.geometry import GeomSpec, build_masks, read_shape
    for mode in ("CUSTOM", "CUSTOM_REF"):
        (tmp_path / "run.toml").write_text(base + f'[geometry]\nGEOM_MODE = "{mode}"\ngeom_file = "tri.toml"\n')
        c = load_config(str(tmp_path / "run.toml"))
        want, _ = build_masks(GeomSpec(GEOM_MODE=mode), c.grid, read_shape(str(tmp_path / "tri.toml")))
        assert np.array_equal(c.masks.f, want.f) and (c.masks.f == 0).sum() > 10
    (tmp_path / "run.toml").write_text(base + '[geometry]\nGEOM_MODE = "CUSTOM"\ngeom_file = "none.toml"\n')
    with pytest.raises(ValueError, match="none.toml"):
        load_config(str(tmp_path / "run.toml"))


def test_material_keys():
    """config.jl:32-90: xi_ratio → m = 1/(R² a/β), g → m = 1/g, per-axis masses, γ as a matrix,
    σn "dirty"."""
    md, info = material({"kappa": 2.0, "gamma": [[0, -0.1, 0.2], [-0.1, 0, 0.0], [0.2, 0.0, 0]],
                         "band": [{"name": "s", "m_x": 4.0}, {"name": "p", "xi_ratio": 2.0, "a": 0.5, "xi_ratio_z": 1.0},
                                  {"g": 0.25, "beta": 2.0, "gamma_kwt": 3.0, "Gamma": 5.0}]})
    b1, b2, b3 = md.bands
    assert b1.ma == (4.0, 1.0, 1.0) and b1.m == 4.0
    assert b2.ma == (1 / (4 * 0.5), 1 / (4 * 0.5), 1 / 0.5)
    assert b3.m == 4.0 and b3.beta == 2.0 and b3.gkwt == 3.0 and b3.Gamma == 5.0
    assert md.gamma[0][1] == -0.1 and md.gamma[0][2] == 0.2 and info["band_names"] == ["s", "p", "band 3"]
    md2, _ = material({"kappa": 2.0, "sigma_n": "dirty", "band": [{"a": 1.0}]})
    assert md2.sigma_n != 1.0
    with pytest.raises(ValueError, match="gamma12 is for two bands"):
        material({"kappa": 2.0, "gamma12": 0.1, "band": [{}, {}, {}]})


def test_scheme_keys_and_refusals(tmp_path):
    txt = TEST_CONFIG_RUN.replace('rkl = false', 'rkl = false\nsigma_vac = 0.01\npcg_tol = 1e-12\nnoise = 1e-3\n'
                                                 'noise_seed = 4\nsubsteps = [1, 0]').replace("nz = 4", "nz = 4\nbuf = 2")
    f = tmp_path / "a.toml"
    f.write_text(txt)
    st = stepper(load_config(str(f)))
    assert st.field_mode == 2 and st.pcg_tol == 1e-12 and st.noise.T == 1e-3 and st.noise.seed == 4
    g = tmp_path / "heat.toml"
    g.write_text(TEST_CONFIG_RUN + "[heat]\nD_T = 0.25\n")
    with pytest.raises(NotImplementedError):
        stepper(load_config(str(g)))
    h = tmp_path / "drag.toml"
    h.write_text(TEST_CONFIG_RUN.replace("rkl = false", "rkl = false\ndrag = 0.05"))
    with pytest.raises(NotImplementedError):
        stepper(load_config(str(h)))


def test_scheme_defaults_of_a_run_folder(tmp_path):
    """config.jl:181-204: the run.toml of a folder that holds a ladder started before the defaults
    changed keeps RKL stages and tol 1e-11; a key of [scheme] wins."""
    d = tmp_path / "old"
    (d / "restart").mkdir(parents=True)
    (d / "run.toml").write_text(TEST_CONFIG_RUN.replace('rkl = false', 'pcg_fast = true'))
    c = load_config(str(d / "run.toml"))
    assert c.scheme["rkl"] is True and c.scheme["pcg_tol"] == 1e-11 and c.scheme["pcg_fast"] is True
    e = tmp_path / "new"
    e.mkdir()
    (e / "run.toml").write_text(TEST_CONFIG_RUN)
    assert "pcg_tol" not in load_config(str(e / "run.toml")).scheme


def test_start_states(tmp_path):
    f = tmp_path / "run.toml"
    f.write_text(TEST_CONFIG_RUN)
    c = load_config(str(f))
    assert start_state(c) is None
    s = initial_state(c.model, c.grid, c.masks, c.H)
    for g in s.G:
        g *= 0.3 + 0.2j
    save_state(str(tmp_path / "seed.npz"), s)
    f2 = tmp_path / "run2.toml"
    f2.write_text(TEST_CONFIG_RUN.replace("H = 0.03", 'H = 0.03\nstart = "seed.npz"'))
    s2 = start_state(load_config(str(f2)))
    assert np.array_equal(s2.G[0], s.G[0])
    f3 = tmp_path / "run3.toml"
    f3.write_text(TEST_CONFIG_RUN.replace("H = 0.03", 'H = 0.03\nstart = "nowhere"'))
    with pytest.raises(ValueError, match="neither"):
        start_state(load_config(str(f3)))
