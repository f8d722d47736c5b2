"""Loading of the golden fixtures (tests/golden/*.npz, converted from the reference's
gtdgl/test/golden/*.jld2) and the release-1.2 stepper they were made with
(gtdgl/test/test_regression12.jl:12-34)."""
from __future__ import annotations

import json
import os

import numpy as np

from juliatdgl_amd.scheme import Band, Grid, Masks, Model, make_stepper, two_band
from juliatdgl_amd.state import State

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = ("k2_R13_explicit", "k2_R4_prod", "k25_R4_prod")


def load_case(name):
    z = np.load(os.path.join(GOLDEN, name + ".npz"))
    par = json.loads(bytes(z["par_json"]).decode())
    ini = {k[4:]: z[k] for k in z.files if k.startswith("ini_")}
    fin = {k[4:]: z[k] for k in z.files if k.startswith("fin_")}
    mk = Masks(*(np.asfortranarray(z[k]) for k in ("fmask", "taux", "tauy", "tauz", "taupsi")))
    return par, ini, fin, mk


def stepper_12(par, mk):
    """_stepper_12 (test_regression12.jl:12-28), without heat and drag."""
    md = two_band(xi_ratio=np.sqrt(1 / (par["m_ratio"] * par["a2n"])), a2n=par["a2n"], gamma12=par["gamma12"],
                  gamma_ratio=par["gamma_ratio"], kappa=par["kappa"])
    b1, b2 = md.bands
    md = Model.make((b1, Band.make(m=par["m_ratio"], Gamma=b2.Gamma, a=b2.a, beta=b2.beta)),
                    [[0.0, par["gamma12"]], [par["gamma12"], 0.0]], kappa=par["kappa"])
    gr = Grid.make(par["Nx_sc"], par["Ny_sc"], par["Nz_sc"], dx=par["dx"], buf=par["buf"])
    sub = (1, 1) if par["pi_substeps_cfg"] == 1 else (1, 0)
    return make_stepper(md, gr, mk, implicit=bool(par["implicit"]), substeps=sub, dt_scale=par["dt_scale"],
                        extrap=bool(par["extrap"]), selffield_bc="slab", rkl=False)


def state_from(gr, d):
    """_state_from (test_regression12.jl:30-34)."""
    return State.from_fields(gr, (d["G1"], d["G2"]), (d["thetax"], d["thetay"], d["thetaz"]),
                             (d["phix"], d["phiy"], d["phiz"]), (d["hx"], d["hy"], d["hz"]))


def same(a, b, tol=1e-10):
    """Every entry to tol of max(1, |b|) (test_regression12.jl:51-57)."""
    a = np.asarray(a)
    b = np.asarray(b)
    assert a.shape == b.shape, (a.shape, b.shape)
    d = np.max(np.abs(a - b) / np.maximum(1.0, np.abs(b))) if a.size else 0.0
    return d <= tol, d
