"""The geometry (juliatdgl_amd/geometry.py) against the masks juliaTDGL 1.2 wrote
(gtdgl/test/golden/masks.jld2, made by make_masks.jl): every catalogue mode and its _REF mirror on
two meshes, the z modes, the shifts, the prism angle, the FGRAD keys and every shape file as
CUSTOM and CUSTOM_REF, to the bit; the box's link and node masks; and what the reference's
test_geometry.jl checks beyond the goldens (the mirror is exact, the rasteriser's special cases)."""
import json
import math
import os

import numpy as np
import pytest

from juliatdgl_amd import geometry as geo
from juliatdgl_amd.scheme import Grid

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _golden():
    z = np.load(os.path.join(GOLDEN, "catalogue_masks.npz"))
    return z, json.loads(bytes(z["meta_json"]).decode())


Z, META = _golden()
CASES = sorted(META["cases"])


def _masks(name):
    m = META["cases"][name]
    ma, g = m["malla"], dict(m["geo"])
    gr = Grid.make(ma["Nx_sc"], ma["Ny_sc"], ma["Nz_sc"], dx=ma["dx"], buf=ma["buf"])
    elems = geo.shape_elements(META["shapes"][g["geom_file"]], g["geom_file"]) if "geom_file" in g else None
    return geo.build_masks(geo.GeomSpec.from_keys(g), gr, elems), gr


def test_every_golden_case_is_there():
    modes = {META["cases"][k]["geo"]["GEOM_MODE"] for k in CASES}
    for m in geo.CATALOGUE:
        assert m in modes and m + "_REF" in modes, m
    assert {"CUSTOM", "CUSTOM_REF"} <= modes
    assert len(CASES) == 128 and len(META["shapes"]) == 23


@pytest.mark.parametrize("name", CASES)
def test_masks_are_release_12s(name):
    (mk, sh), _ = _masks(name)
    assert mk.f.shape == Z[name].shape
    assert np.array_equal(mk.f, Z[name]), (name, int((mk.f != Z[name]).sum()))
    assert sh.vac is None and sh.mat is None


def test_box_link_and_node_masks():
    (mk, _), gr = _masks("WEDGE_40x24x6")
    for t, a in (("taux", mk.tx), ("tauy", mk.ty), ("tauz", mk.tz), ("taupsi", mk.tpsi)):
        assert np.array_equal(a, Z["_box_40x24x6_" + t].astype(float)), t


@pytest.mark.parametrize("mode", [m for m in geo.CATALOGUE if m not in ("SIM",)])
def test_ref_is_the_exact_mirror_of_the_interior_nodes(mode):
    """mirror_defects! maps the interior nodes y1+1..y2 onto themselves (geometry.jl:177-192): the
    _REF mask on them is the mode's reflected j -> (y1 + y2 + 1) - j (FGRAD's sawtooth reversed)."""
    gr = Grid.make(40, 24, 6, dx=0.5, buf=1)
    a, _ = geo.build_masks(geo.GeomSpec(GEOM_MODE=mode), gr)
    b, _ = geo.build_masks(geo.GeomSpec(GEOM_MODE=mode + "_REF"), gr)
    x1, x2, y1, y2, z1, z2 = gr.x1, gr.x2, gr.y1, gr.y2, gr.z1, gr.z2
    for j in range(y1 + 1, y2 + 1):
        jm = (y1 + y2 + 1) - j
        fa, fb = a.f[x1:x2, j - 1, z1 - 1:z2], b.f[x1:x2, jm - 1, z1 - 1:z2]
        if mode == "FGRAD":  # t and 1 - (1 - t) round apart
            assert np.allclose(fa, fb, rtol=0, atol=1e-15), (mode, j)
        else:
            assert np.array_equal(fa, fb), (mode, j)


def test_rasteriser_special_cases():
    # ties away from zero, as RoundNearestTiesAway
    assert [geo.oround(x) for x in (0.5, 1.5, 2.5, -0.5, -2.5, 2.4999999999999996)] == [1, 2, 3, -1, -3, 2]
    # the triangle n = 3, R = 6 keeps its 11-cell base (the +1e-9 of in_ngon)
    jj, ii = np.meshgrid(np.arange(-8, 9), np.arange(-8, 9), indexing="ij")
    m = geo.in_ngon(ii.astype(float), jj.astype(float), 0, 0, 6.0, 3, math.pi / 2)
    base = jj[m].min()
    assert int(m[jj == base].sum()) == 11
    # a point on a polygon's edge is inside; outside reverses a figure
    assert geo.in_polygon(np.array([0.5]), np.array([0.0]), (0.0, 1.0, 0.0), (0.0, 0.0, 1.0))[0]
    gr = Grid.make(20, 12, 4, dx=0.5, buf=1)
    mk, _ = geo.build_masks(geo.GeomSpec(GEOM_MODE="CUSTOM"), gr,
                            geo.shape_elements({"shape": [{"type": "circle", "x": 5.0, "y": 3.0, "r": 2.0, "f": 0.2,
                                                           "outside": True}]}))
    x1, y1, z1 = gr.x1, gr.y1, gr.z1
    assert mk.f[x1 - 1 + 10, y1 - 1 + 6, z1 - 1] == 1.0 and mk.f[x1 - 1, y1 - 1, z1 - 1] == 0.2


def test_old_names_and_errors():
    e = geo.shape_elements({"forma": [{"tipo": "circulo", "x": 1.0, "y": 1.0, "r": 1.0, "z": "todo"}]})
    assert e[0]["type"] == "circle" and e[0]["z"] == "all"
    with pytest.raises(ValueError, match="unknown type"):
        geo.shape_elements({"shape": [{"type": "blob"}]})
    with pytest.raises(ValueError, match="unknown GEOM_MODE"):
        geo.build_masks(geo.GeomSpec(GEOM_MODE="BLOB"), Grid.make(8, 6, 4, dx=0.5))
    with pytest.raises(ValueError, match="z_mode"):
        geo.build_masks(geo.GeomSpec(GEOM_MODE="ESFERA", z_mode="sideways"), Grid.make(40, 24, 6, dx=0.5))


def test_vacuum_sculpts_the_sample():
    """sculpt! (geometry.jl:537-551): a vacuum region leaves the block (f = τψ = 0, the links to it
    cut, its face ghosts vacuum too); a later f gives the material back; no material left is an
    error; conducting_nodes drops the vacuum from the bar."""
    gr = Grid.make(20, 12, 4, dx=0.5, buf=1)
    el = geo.shape_elements({"shape": [{"type": "rect", "x": 0.0, "y": 0.0, "w": 4.0, "h": 4.0, "vacuum": True},
                                       {"type": "circle", "x": 1.0, "y": 1.0, "r": 0.3, "f": 0.5}]})
    mk, sh = geo.build_masks(geo.GeomSpec(GEOM_MODE="CUSTOM"), gr, el)
    v = sh.vac
    assert mk.vac is v
    x1, y1, z1 = gr.x1, gr.y1, gr.z1
    c = (x1 - 1 + 3, y1 - 1 + 2, z1)  # a vacuum node (X = 1.5, Y = 1 ξσ)
    assert v[c] and mk.f[c] == 0 and mk.tpsi[c] == 0 and mk.tx[c] == 0 and mk.tx[c[0] - 1, c[1], c[2]] == 0
    assert mk.ty[c[0], c[1] - 1, c[2]] == 0 and mk.tz[c[0], c[1], c[2] - 1] == 0
    k = (x1 - 1 + 2, y1 - 1 + 2, z1)  # X = Y = 1: the circle gave the material back
    assert not v[k] and mk.f[k] == 0.5 and mk.tpsi[k] == 1
    assert v[x1 - 1, y1 - 1 + 3, z1]  # the face plane x1 copies the plane after it
    with pytest.raises(NotImplementedError, match="sculpted"):
        from juliatdgl_amd.scheme import make_stepper, two_band
        make_stepper(two_band(xi_ratio=4.0, kappa=2.0), gr, mk, sigma_vac=0.01)
    cond = geo.conducting_nodes(gr, v)
    assert not cond[c] and cond[gr.x2, gr.y2, z1]
    with pytest.raises(ValueError, match="no material"):
        geo.build_masks(geo.GeomSpec(GEOM_MODE="CUSTOM"), gr,
                        geo.shape_elements({"shape": [{"type": "rect", "x": 5.0, "y": 3.0, "w": 40.0, "h": 40.0,
                                                       "vacuum": True}]}))
