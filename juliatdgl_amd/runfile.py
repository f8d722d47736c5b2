"""A run from one TOML file (gtdgl/src/config.jl:1-242), for the options this stepper runs.

    c = load_config("run.toml")        # (model, info, grid, masks, geom, scheme, protocol, J, H, out_dir, raw)
    st = stepper(c)                    # Stepper(c.model, c.grid, c.masks; c.scheme...)
    s = start_state(c)                 # the state the ladders start from ([run] start)

The tables and keys are the reference's:

    [material]  file = "<material>.toml" (merged under the keys given), or inline: kappa, sigma_n
                (a number or "dirty"), T, gamma12 | gamma; [[material.band]] m | g | xi_ratio, Gamma, a,
                beta, gamma_kwt, name; m_x, m_y, m_z or xi_ratio_x, ... for an anisotropic band
    [mesh]      nx, ny, nz, dx, buf (bufx, bufy, bufz)
    [geometry]  the GeomSpec keys: GEOM_MODE (the catalogue SIM, ESFERA, ARRAY, WEDGE, WEDGE_BOTH,
                ESF_CUNA, FGRAD, SAWROW, SPHROW, SAWSPH, STAGGER, FUNNEL, FUNNEL_IN, TRI_Y1, TRI_Y1_INV,
                RECT_Y1, HEX, HEX_STAG and their _REF mirrors, or CUSTOM / CUSTOM_REF with geom_file, a
                shape file next to the run file or in TDGL_SHAPES), wedge_depth, wedge_half, wedge_depth2,
                fgrad_*, x_shift, y_shift, z_mode, prism_phi, inj_margin, ...
    [scheme]    implicit, dt_scale, substeps, extrap, hard, rkl, rotframe, rkl_margin, selffield_bc,
                open_every, sigma_vac, qp_exchange, qp_split, noise, noise_seed, noise_field, pcg_tol,
                pcg_fast, pcg_f32 (threads is the CPU step's and is ignored)
    [protocol]  the Protocol keys
    [run]       J, H, out_dir, start ("meissner", "field_cooled" or a checkpoint file)

A material `file` is looked for as given, next to the run file, and in the directories of
TDGL_MATERIALS (the reference's gtdgl/materials, say); the material of the reference's validation
run files is built in.  Options this stepper does not run (heat, drag, materials per region, a
sculpted sample, a current along y) raise NotImplementedError, as make_stepper does.
"""
from __future__ import annotations

import dataclasses
import math
import os
from dataclasses import dataclass
from typing import Optional

import numpy as np

try:
    import tomllib as _toml
except ModuleNotFoundError:  # Python < 3.11
    import tomli as _toml

from .meter import Protocol
from .scheme import Band, Grid, Masks, Model, Stepper, make_stepper
from .state import State, initial_state

# the material of the reference's validation run files (gtdgl/materials/mgb2_model_II.toml): the
# two-band model of the diode studies, as its keys
BUILTIN_MATERIALS = {
    "mgb2_model_II.toml": {
        "name": "MgB2-like two-band model of the diode studies",
        "kappa": 25.0, "sigma_n": 1.0, "T": 0.0, "gamma12": 1e-3,
        "band": [{"name": "sigma", "m": 1.0, "Gamma": 1.0, "a": 1.0},
                 {"name": "pi", "xi_ratio": 4.0, "Gamma": 1.0, "a": 0.59}],
    },
}

# the defaults of the scheme since release 2.0's change of 26-oct-2026, and before it: a run folder
# that already holds a ladder keeps the ones it started with (config.jl:181-204)
SCHEME_DEFAULTS = {"rkl": "rkc", "pcg_tol": 1e-8, "pcg_fast": True, "pcg_f32": True}
SCHEME_DEFAULTS_BEFORE = {"rkl": True, "pcg_tol": 1e-11, "pcg_fast": False, "pcg_f32": False}

SCHEME_KEYS = {"implicit", "dt_scale", "substeps", "extrap", "hard", "rkl", "rotframe", "rkl_margin", "rkc_damping",
               "selffield_bc", "open_every", "sigma_vac", "qp_exchange", "qp_split", "noise", "noise_seed",
               "noise_field", "pcg_tol", "pcg_fast", "pcg_f32", "pcg_extrap", "pcg_maxit", "dt"}


def _band(d: dict) -> Band:
    """_band (config.jl:32-51): m from xi_ratio as 1/(R² a/β), the axes of an anisotropic band."""
    k = dict(d)
    a = float(k.get("a", 1.0))
    beta = float(k.get("beta", 1.0))
    if "xi_ratio" in k:
        k["m"] = 1 / (float(k["xi_ratio"]) ** 2 * (a / beta))
    m0 = 1 / float(k["g"]) if "g" in k else float(k.get("m", 1.0))
    ax = []
    for c in "xyz":
        if f"m_{c}" in k:
            ax.append(float(k[f"m_{c}"]))
        elif f"xi_ratio_{c}" in k:
            ax.append(1 / (float(k[f"xi_ratio_{c}"]) ** 2 * (a / beta)))
        else:
            ax.append(None)
    mx = None if all(x is None for x in ax) else tuple(m0 if x is None else x for x in ax)
    return Band.make(m=float(k.get("m", 1.0)), g=float(k["g"]) if "g" in k else None, Gamma=float(k.get("Gamma", 1.0)),
                     a=a, beta=beta, gkwt=float(k.get("gamma_kwt", 0.0)), m_axes=mx)


def _find_material(f: str, base_dir: str) -> dict:
    cands = [f] if os.path.isabs(f) else [f, os.path.join(base_dir, f)]
    for d in filter(None, os.environ.get("TDGL_MATERIALS", "").split(os.pathsep)):
        cands.append(os.path.join(d, f))
    for p in cands:
        if os.path.isfile(p):
            with open(p, "rb") as fh:
                return _toml.load(fh)
    if f in BUILTIN_MATERIALS:
        return dict(BUILTIN_MATERIALS[f])
    raise FileNotFoundError(f"material file {f!r}: not found (as given, next to the run file, in TDGL_MATERIALS)")


def _material_dict(d: dict, base_dir: str) -> dict:
    """_material_dict (config.jl:54-59): the file's table with the keys given on top."""
    if "file" not in d:
        return d
    out = _find_material(d["file"], base_dir)
    out.update({k: v for k, v in d.items() if k != "file"})
    return out


def material(d: dict, base_dir: str = "."):
    """material(d) -> (Model, info) (config.jl:67-90)."""
    d = _material_dict(d, base_dir)
    bands = d["band"]
    N = len(bands)
    bs = [_band(b) for b in bands]
    g = np.zeros((N, N))
    if "gamma" in d:
        g[...] = np.array([[float(x) for x in r] for r in d["gamma"]])
    elif "gamma12" in d:
        if N != 2:
            raise ValueError("material: gamma12 is for two bands; give the matrix `gamma`")
        g[0, 1] = g[1, 0] = float(d["gamma12"])
    sn = d.get("sigma_n", 1.0)
    sn = "dirty" if sn == "dirty" else float(sn)
    md = Model.make(bs, g, kappa=float(d["kappa"]), sigma_n=sn, T=float(d.get("T", 0.0)))
    info = {"name": d.get("name", ""), "regime": d.get("regime", ""), "sources": d.get("sources", {}),
            "band_names": [b.get("name", f"band {mu + 1}") for mu, b in enumerate(bands)]}
    return md, info


def shape_path(f: str, base_dir: str) -> str:
    """A [geometry] geom_file (config.jl:182-186): as given when absolute, else next to the run file,
    else in the directories of TDGL_SHAPES (the reference's gtdgl/shapes, say)."""
    if os.path.isabs(f):
        return f
    cands = [os.path.join(base_dir, f)] + [os.path.join(d, f) for d in os.environ.get("TDGL_SHAPES", "").split(os.pathsep)
                                           if d]
    for c in cands:
        if os.path.isfile(c):
            return c
    raise ValueError(f"geom_file {f!r}: not next to the run file nor in TDGL_SHAPES")


def geometry_masks(gk: dict, gr: Grid, base_dir: str = ".") -> Masks:
    """build_masks(GeomSpec(; gk...), gr) (geometry.jl:485-527; juliatdgl_amd.geometry): every
    catalogue mode and its _REF mirror, the shifts, z_mode, prism_phi and the shape files (CUSTOM,
    CUSTOM_REF).  The current runs along x (current_axis = "x", the stepper's injection)."""
    from .geometry import GeomSpec, build_masks
    if str(gk.get("current_axis", "x")).lower() != "x":
        raise NotImplementedError("[geometry] current_axis other than x is not run here (the current is injected along x)")
    g = dict(gk)
    if g.get("geom_file"):
        g["geom_file"] = shape_path(str(g["geom_file"]), base_dir)
    mk, _ = build_masks(GeomSpec.from_keys(g), gr)
    return mk


def _holds_ladder(od: str) -> bool:
    """_holds_ladder (config.jl:185-186)."""
    out = os.path.join(od, "output")
    return (os.path.isfile(os.path.join(od, "j_history.csv")) or os.path.isdir(os.path.join(od, "restart")) or
            (os.path.isdir(out) and any(f.startswith("rungs_") for f in os.listdir(out))))


def _scheme_defaults(sk: dict, od: str) -> None:
    """_scheme_defaults! (config.jl:187-194)."""
    f = os.path.join(od, "scheme_defaults.toml")
    if os.path.isfile(f):
        with open(f, "rb") as fh:
            d = _toml.load(fh)
    else:
        d = SCHEME_DEFAULTS_BEFORE if _holds_ladder(od) else SCHEME_DEFAULTS
    for k, v in d.items():
        sk.setdefault(k, v)
    if os.path.isfile(f):
        for k, v in SCHEME_DEFAULTS_BEFORE.items():
            sk.setdefault(k, v)


@dataclass
class RunConfig:
    """load_config's NamedTuple (config.jl:209-242); `scheme` holds make_stepper's keywords."""
    model: Model
    info: dict
    grid: Grid
    masks: Masks
    geom: dict
    scheme: dict
    protocol: Protocol
    J: list
    H: float
    out_dir: str
    raw: dict
    path: str


def load_config(path: str) -> RunConfig:
    """load_config(path) (config.jl:209-242)."""
    with open(path, "rb") as fh:
        raw = _toml.load(fh)
    base = os.path.dirname(os.path.abspath(path))
    md, info = material(raw["material"], base)
    ms = raw["mesh"]
    b = ms.get("buf", 1)
    gr = Grid.make(int(ms["nx"]), int(ms["ny"]), int(ms["nz"]), dx=float(ms["dx"]), buf=b,
                   bufx=ms.get("bufx", b), bufy=ms.get("bufy", b), bufz=ms.get("bufz", b))
    if "materials" in raw:
        raise NotImplementedError("[materials.<name>]: materials per region are not run here")
    gk = dict(raw.get("geometry", {}))
    mk = geometry_masks(gk, gr, base)
    if ms.get("ghost_links", True) is False:
        raise NotImplementedError("[mesh] ghost_links = false is not run here")
    sk = dict(raw.get("scheme", {}))
    if os.path.basename(path) == "run.toml" and _holds_ladder(base):
        _scheme_defaults(sk, base)
    sk.pop("threads", None)
    if "heat" in raw:
        sk["heat"] = dict(raw["heat"])
    if "substeps" in sk:
        sk["substeps"] = tuple(int(x) for x in sk["substeps"])
    if "dt_scale" in sk:
        sk["dt_scale"] = float(sk["dt_scale"])
    if "sigma_vac" in sk:
        sk["sigma_vac"] = float(sk["sigma_vac"])
    if "qp_split" in sk:
        sk["qp_split"] = tuple(float(x) for x in sk["qp_split"])
    if isinstance(sk.get("qp_exchange"), list):
        sk["qp_exchange"] = np.array([[float(x) for x in r] for r in sk["qp_exchange"]])
    if isinstance(sk.get("rkl"), str):
        sk["rkl"] = sk["rkl"].lower()
    pk = dict(raw.get("protocol", {}))
    names = {f.name for f in dataclasses.fields(Protocol)}
    pr = Protocol(**{k: (float(v) if isinstance(getattr(Protocol, k, None), float) and isinstance(v, int) else v)
                     for k, v in pk.items() if k in names})
    rn = raw["run"]
    J = [float(x) for x in rn["J"]]
    return RunConfig(md, info, gr, mk, gk, sk, pr, J, float(rn.get("H", 0.0)), rn.get("out_dir", ""), raw,
                     os.path.abspath(path))


def stepper(c: RunConfig) -> Stepper:
    """Stepper(c.model, c.grid, c.masks; c.scheme...)."""
    return make_stepper(c.model, c.grid, c.masks, **c.scheme)


def start_state(c: RunConfig, backend=None) -> Optional[State]:
    """start_state(c, path) (config.jl:291-318): None for "meissner" (the Meissner seed), the state
    of a checkpoint file (relative to the run file), or for "field_cooled" the sample held normal at
    the bath temperature start_T_hot for start_t_hot, then Ψ = 10⁻³ × the noise of start_seed on the
    conductor (numpy's PCG64 stream: equal in distribution to the reference's Xoshiro draws).  The
    hot steps run on `backend(stepper) -> device stepper` (juliatdgl_amd.hip.HipStepper)."""
    rn = c.raw["run"]
    how = rn.get("start", "meissner")
    if how == "meissner":
        return None
    if how == "field_cooled":
        if backend is None:
            raise ValueError("start = 'field_cooled' needs a backend to run the hot steps on")
        md = c.model
        hot = Model.make(md.bands, np.array(md.gamma), kappa=md.kappa, sigma_n=md.sigma_n,
                         T=float(rn.get("start_T_hot", 1.5)))
        sh = make_stepper(hot, c.grid, c.masks, **c.scheme)
        s = initial_state(md, c.grid, c.masks, c.H)
        dev = backend(sh)
        dev.upload(s)
        dev.step(round(float(rn.get("start_t_hot", 150.0)) / sh.dt), c.H, 0.0)
        s = dev.download()
        rng = np.random.default_rng(int(rn.get("start_seed", 1)))
        for mu in range(md.N):
            n = rng.standard_normal((2,) + s.G[mu].shape)
            s.G[mu] = np.asfortranarray(c.masks.tpsi * (1e-3 * (n[0] + 1j * n[1])))
            s.F[mu] = s.G[mu].copy(order="F")
        return s
    from .checkpoint import load_state
    p = how if os.path.isabs(how) else os.path.join(os.path.dirname(c.path), how)
    if not os.path.isfile(p):
        raise ValueError(f"start = {how!r} in [run]: neither 'meissner', 'field_cooled' nor a state file")
    return load_state(p, c.grid)[0]
