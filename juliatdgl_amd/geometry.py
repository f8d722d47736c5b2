"""The sample's shape: the defect catalogue of GEOM_MODE and the shape files of GEOM_MODE = CUSTOM /
CUSTOM_REF, as the masks f, τx, τy, τz, τψ (and, for a sculpted sample, its vacuum; for a shape
with materials, the material index of every node).

Restates gtdgl/src/geometry.jl (GeomSpec :22-38, the rasteriser paint_defect! :92-158, the
transformations :163-238, the catalogue build_defect_list :268-476, build_masks :485-527,
sculpt! :537-551, conducting_nodes :559-570) and gtdgl/src/shapes.jl (the shape-file format,
read_shape :84-96, _dentro :142-186, _pixel :189-208, _voxel :216-235, rasterize_shape! :245-287):
the same node sets bit for bit, checked against every mask of gtdgl/test/golden/masks.jld2 (every
catalogue mode and its mirror on two meshes, the z modes, the shifts, the prism angle, every shape
file).  Indices are the reference's (1-based x1..x2, y1..y2, z1..z2 of `Grid`); the arrays are
node-shaped and 0-based.  Host code, vectorised over the nodes: one pass per defect or shape
element, in the order the reference paints them (the last writer of a node wins either way).
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field
from fractions import Fraction
from typing import List, Optional

import numpy as np

TOL = 1e-9


def oround(x) -> int:
    """round(Int, x, RoundNearestTiesAway) of a float or a Fraction."""
    if isinstance(x, Fraction):
        a = abs(x)
        r = math.floor(a + Fraction(1, 2))
        return int(r if x >= 0 else -r)
    a = abs(float(x))
    r = math.floor(a)
    r += 1 if a - r >= 0.5 else 0
    return int(r if x >= 0 else -r)


@dataclass
class GeomSpec:
    """GeomSpec (geometry.jl:22-38): the mode, the shape file and the catalogue's parameters, with
    the reference's names and defaults."""
    GEOM_MODE: str = "SIM"
    geom_file: str = ""
    wedge_depth: float = 0.35
    wedge_half: float = 0.20
    wedge_depth2: float = 0.0
    fgrad_flo: float = 1.0
    fgrad_fhi: float = 1.5
    fgrad_nper: int = 2
    fgrad_dir: int = 1
    inj_margin: int = 6
    strict_margin: bool = True
    x_shift: float = 0.0
    y_shift: float = 0.0
    z_mode: str = "through"
    prism_phi: float = float("nan")
    current_axis: str = "x"
    verbose: bool = False
    materials: List[str] = field(default_factory=list)

    @classmethod
    def from_keys(cls, d: dict) -> "GeomSpec":
        kw = {}
        for k, v in d.items():
            if k not in cls.__dataclass_fields__:
                raise ValueError(f"[geometry] unknown key {k!r}")
            kw[k] = list(v) if k == "materials" else v
        return cls(**kw)


@dataclass
class Defect:
    """A catalogue defect (geometry.jl:41-65); the fields a type does not use stay None."""
    type: str
    xc: Optional[int] = None
    yc: Optional[int] = None
    zc: Optional[int] = None
    R: Optional[float] = None
    n: Optional[int] = None
    phi: Optional[float] = None
    hx: Optional[int] = None
    hy: Optional[int] = None
    hz: Optional[int] = None
    half: Optional[int] = None
    depth: Optional[int] = None
    f: float = 1.0
    xa: Optional[int] = None
    xb: Optional[int] = None
    ya: Optional[int] = None
    yb: Optional[int] = None
    flo: Optional[float] = None
    fhi: Optional[float] = None
    nper: Optional[int] = None
    dir: Optional[int] = None
    zmode: Optional[str] = None
    zlo: Optional[int] = None
    zhi: Optional[int] = None


# ---------------------------------------------------------------------------- the rasteriser
def in_ngon(px, py, cx, cy, R, n, phi):
    """Inside the regular n-gon of circumradius R turned by phi (geometry.jl:79-90), element-wise:
    inside the inscribed circle, outside the circumscribed one, else against the edge's radius
    a / cos(θ mod 2π/n − π/n); every comparison with the reference's 1e-9."""
    dx = px - cx
    dy = py - cy
    r = np.hypot(dx, dy)
    a = R * math.cos(math.pi / n)
    th = np.arctan2(dy, dx) - phi
    seg = np.mod(th, 2 * math.pi / n) - math.pi / n
    return (r <= a + 1e-9) | ((r <= R + 1e-9) & (r <= a / np.cos(seg) + 1e-9))


def _block(x1, x2, y1, y2, zA, zB):
    """Index grids (1-based, as the reference's) of the nodes x1:x2, y1:y2, zA:zB."""
    return np.meshgrid(np.arange(x1, x2 + 1), np.arange(y1, y2 + 1), np.arange(zA, zB + 1), indexing="ij")


def paint_defect(fmask, d: Defect, x1, x2, y1, y2, z1, z2):
    """paint_defect! (geometry.jl:92-158): f = d.f on the nodes of x1:x2, y1:y2 and the defect's z
    range that the defect covers."""
    zA = z1 if d.zlo is None else max(z1, d.zlo)
    zB = z2 if d.zhi is None else min(z2, d.zhi)
    if d.type == "fgrad":
        # the sawtooth in y: t over [ya, yb] (reversed for dir < 0), phase mod(t nper, 1)
        Ly = max(1, d.yb - d.ya)
        xa, xb = max(x1, d.xa), min(x2, d.xb)
        for j in range(max(y1, d.ya), min(y2, d.yb) + 1):
            t = (j - d.ya) / Ly
            if d.dir < 0:
                t = 1 - t
            fase = math.fmod(t * d.nper, 1.0)
            if fase < 0:
                fase += 1.0
            fj = d.fhi + (d.flo - d.fhi) * fase
            if xa <= xb and zA <= zB:
                fmask[xa - 1:xb, j - 1, zA - 1:zB] = fj
        return fmask
    if zA > zB:
        return fmask
    I, J, K = _block(x1, x2, y1, y2, zA, zB)
    if d.type == "sphere":
        m = (I - d.xc) ** 2 + (J - d.yc) ** 2 + (K - d.zc) ** 2 <= d.R ** 2
    elif d.type == "cylinder":
        m = (I - d.xc) ** 2 + (J - d.yc) ** 2 <= d.R ** 2
    elif d.type == "box":
        m = (np.abs(I - d.xc) <= d.hx) & (np.abs(J - d.yc) <= d.hy) & (np.abs(K - d.zc) <= d.hz)
    elif d.type == "prism":
        m = in_ngon(I.astype(float), J.astype(float), d.xc, d.yc, d.R, d.n, d.phi)
    elif d.type == "wedge":
        # the notch grows from its face along dir: at depth `layer` the steep wall spans
        # max(1, round(2 frac)) cells before xc and the gentle ramp round(half frac) after
        dd = 1 if d.dir is None else d.dir
        m = np.zeros(I.shape, dtype=bool)
        for j in range(y1, y2 + 1):
            layer = dd * (j - d.yc)
            if 0 <= layer < d.depth:
                frac = 1 - layer / d.depth
                w_steep = max(1, oround(2 * frac))
                w_gentle = oround(d.half * frac)
                dxr = np.arange(x1, x2 + 1) - d.xc
                m[:, j - y1, :] = ((dxr >= -w_steep) & (dxr <= w_gentle))[:, None]
    else:
        raise ValueError(f"unknown defect type: {d.type}")
    sub = fmask[x1 - 1:x2, y1 - 1:y2, zA - 1:zB]
    sub[m] = d.f
    return fmask


# ---------------------------------------------------------------------------- transformations
def shift_defects(defects, ax: str, dcell: int):
    """shift_defects! (geometry.jl:163-175)."""
    for d in defects:
        if ax.lower() == "x":
            if d.xc is not None:
                d.xc += dcell
            if d.xa is not None:
                d.xa += dcell
                d.xb += dcell
        else:
            if d.yc is not None:
                d.yc += dcell
            if d.ya is not None:
                d.ya += dcell
                d.yb += dcell
    return defects


def mirror_defects(defects, cfg: GeomSpec, x1, x2, y1, y2):
    """mirror_defects! (geometry.jl:193-217): the exact mirror of the interior nodes along the
    vortex motion, y -> (y1 + y2 + 1) - y for J along x (x -> (x1 + x2 + 1) - x otherwise)."""
    mirror_y = cfg.current_axis.lower() == "x"
    for d in defects:
        if d.type == "wedge" and not mirror_y:
            raise ValueError("mirror_defects: mirroring 'wedge' in x is not defined (use J||x)")
        if mirror_y:
            if d.yc is not None:
                d.yc = (y1 + y2 + 1) - d.yc
            if d.phi is not None:
                d.phi = -d.phi
            if d.type == "fgrad":
                ya0, yb0 = d.ya, d.yb
                d.ya = (y1 + y2 + 1) - yb0
                d.yb = (y1 + y2 + 1) - ya0
                d.dir = -d.dir
            if d.type == "wedge":
                d.dir = -d.dir
        else:
            if d.xc is not None:
                d.xc = (x1 + x2 + 1) - d.xc
            if d.phi is not None:
                d.phi = math.pi - d.phi
            if d.type == "fgrad":
                raise ValueError("mirror_defects: fgrad is only defined for J||x")
    return defects


def apply_zmode(defects, cfg: GeomSpec, z1, z2):
    """apply_zmode! (geometry.jl:220-238): through, half (z1 .. the middle) or buried (a quarter
    of the thickness, at least one plane, kept from both faces)."""
    zc = oround((z1 + z2) / 2)
    for d in defects:
        zm = (cfg.z_mode if d.zmode is None else d.zmode).lower()
        if zm == "through":
            d.zlo, d.zhi = z1, z2
        elif zm == "half":
            d.zlo, d.zhi = z1, zc
        elif zm == "buried":
            q = max(1, oround((z2 - z1) / 4))
            d.zlo, d.zhi = z1 + q, z2 - q
        else:
            raise ValueError(f"unknown z_mode: {zm} (use through|half|buried)")
    return defects


def _linspace(a, b, n):
    """range(a, b, length=n) of the reference, exact: the node values as Fractions."""
    a, b = Fraction(a), Fraction(b)
    return [a + (b - a) * i / (n - 1) for i in range(n)] if n > 1 else [a]


def _phi_cat(cfg: GeomSpec, default: float) -> float:
    return default if math.isnan(cfg.prism_phi) else cfg.prism_phi


CATALOGUE = ("SIM", "ESFERA", "ARRAY", "WEDGE", "WEDGE_BOTH", "ESF_CUNA", "FGRAD", "SAWROW", "SPHROW", "SAWSPH",
             "STAGGER", "FUNNEL", "FUNNEL_IN", "TRI_Y1", "TRI_Y1_INV", "RECT_Y1", "HEX", "HEX_STAG")


def build_defect_list(cfg: GeomSpec, x1, x2, y1, y2, z1, z2) -> List[Defect]:
    """build_defect_list (geometry.jl:268-476): the defects of a catalogue mode (J along x), then
    the shifts, the _REF mirror and the z ranges."""
    yc = oround((y1 + y2) / 2)
    zc = oround((z1 + z2) / 2)
    xc = oround((x1 + x2) / 2)
    nx_sc, ny_sc = x2 - x1 + 1, y2 - y1 + 1
    D: List[Defect] = []
    mode = cfg.GEOM_MODE.upper()
    is_ref = len(mode) > 4 and mode.endswith("_REF")
    if is_ref:
        mode = mode[:-4]
    mg = cfg.inj_margin
    xsafe_lo, xsafe_hi = (x1 + mg, x2 - mg) if cfg.current_axis.lower() == "x" else (x1, x2)
    if mode == "SIM":
        pass
    elif mode == "ESFERA":
        D.append(Defect("sphere", xc=xc, yc=yc, zc=zc, R=6.0, f=0.0))
    elif mode == "ARRAY":
        xs = [oround(v * nx_sc) + x1 for v in (0.30, 0.50, 0.70)]
        for s, R in zip(xs, (4.0, 6.0, 8.0)):
            D.append(Defect("sphere", xc=s, yc=yc, zc=zc, R=R, f=0.0))
    elif mode == "WEDGE":
        D.append(Defect("wedge", xc=xc, yc=y1, dir=1, half=oround(cfg.wedge_half * nx_sc),
                        depth=oround(cfg.wedge_depth * ny_sc), f=0.0))
    elif mode == "WEDGE_BOTH":
        D.append(Defect("wedge", xc=xc, yc=y1, dir=1, half=oround(cfg.wedge_half * nx_sc),
                        depth=oround(cfg.wedge_depth * ny_sc), f=0.0))
        if cfg.wedge_depth2 > 0:
            D.append(Defect("wedge", xc=xc, yc=y2 + 1, dir=-1, half=oround(cfg.wedge_half * nx_sc),
                            depth=oround(cfg.wedge_depth2 * ny_sc), f=0.0))
    elif mode == "ESF_CUNA":
        D.append(Defect("sphere", xc=xc, yc=yc, zc=zc, R=6.0, f=0.0))
        D.append(Defect("wedge", xc=xc + oround(0.18 * nx_sc), yc=y1, dir=1, half=oround(0.10 * nx_sc),
                        depth=oround(0.22 * ny_sc), f=0.0))
    elif mode == "FGRAD":
        D.append(Defect("fgrad", xa=xsafe_lo, xb=xsafe_hi, ya=y1, yb=y2, flo=cfg.fgrad_flo, fhi=cfg.fgrad_fhi,
                        nper=cfg.fgrad_nper, dir=cfg.fgrad_dir, f=float("nan")))
    elif mode in ("SAWROW", "SAWSPH"):
        nteeth = 5
        dep = oround(0.55 * ny_sc)
        xlo, xhi = x1 + mg + 2, x2 - mg - 2
        half = max(3, (xhi - xlo) // (2 * nteeth) - 1)
        for s in _linspace(xlo + half, xhi - half, nteeth):
            D.append(Defect("wedge", xc=oround(s), yc=y1, dir=1, half=half, depth=dep, f=0.0))
        if mode == "SAWSPH":
            zc0 = oround((z1 + z2) / 2)
            ysph = y2 - oround(0.25 * ny_sc)
            xsph = [oround(v * (x2 - x1)) + x1 for v in (0.30, 0.50, 0.70)]
            for s, R in zip(xsph, (3.0, 5.0, 4.0)):
                D.append(Defect("sphere", xc=s, yc=ysph, zc=zc0, R=R, f=0.0))
    elif mode == "SPHROW":
        zc0, yc0 = oround((z1 + z2) / 2), oround((y1 + y2) / 2)
        xs = [oround(v * (x2 - x1)) + x1 for v in (0.30, 0.50, 0.70)]
        for s, R in zip(xs, (3.0, 6.0, 4.0)):
            D.append(Defect("sphere", xc=s, yc=yc0, zc=zc0, R=R, f=0.0))
    elif mode == "STAGGER":
        Rtri = 5.0
        xs = [oround(v) for v in _linspace(xsafe_lo + 5, xsafe_hi - 5, 4)]
        dyr = oround(0.22 * ny_sc)
        for s_, xv in enumerate(xs, start=1):
            ys = [yc - dyr, yc + dyr] if s_ % 2 == 1 else [yc - 2 * dyr, yc, yc + 2 * dyr]
            for yv in ys:
                if y1 + 2 < yv < y2 - 2:
                    D.append(Defect("prism", xc=xv, yc=oround(yv), zc=zc, R=Rtri, n=3, phi=_phi_cat(cfg, 0.0), f=0.0))
    elif mode in ("FUNNEL", "FUNNEL_IN"):
        Rtri, dyrow, per = 6.0, 10, 16
        inset = 4 if mode == "FUNNEL_IN" else 0
        xa, xb = xsafe_lo + Rtri, xsafe_hi - Rtri
        step = per / 2
        xs = []
        v = xa
        while v <= xb:
            xs.append(v)
            v = xa + step * len(xs)
        if not xs:
            raise ValueError("FUNNEL: the mesh is too short in x to place triangles inside the margin")
        for s, xv in enumerate(xs, start=1):
            yrow = yc - dyrow / 2 if s % 2 == 1 else yc + dyrow / 2
            if inset > 0:
                ymax = y2 - Rtri - inset
                ymin = y1 + Rtri * math.cos(math.pi / 3) + inset
                yrow = min(max(float(yrow), ymin), ymax)
            D.append(Defect("prism", xc=oround(xv), yc=oround(yrow), zc=zc, R=Rtri, n=3,
                            phi=_phi_cat(cfg, math.pi / 2), f=0.0))
    elif mode == "TRI_Y1":
        Rtri = 6.0
        D.append(Defect("prism", xc=xc, yc=y1 + oround(Rtri * math.cos(math.pi / 3)), zc=zc, R=Rtri, n=3,
                        phi=_phi_cat(cfg, math.pi / 2), f=0.0))
    elif mode == "TRI_Y1_INV":
        Rtri = 6.0
        D.append(Defect("prism", xc=xc, yc=y1 + oround(Rtri), zc=zc, R=Rtri, n=3, phi=_phi_cat(cfg, -math.pi / 2), f=0.0))
    elif mode == "RECT_Y1":
        hx, hy = 5, 2
        D.append(Defect("box", xc=xc, yc=y1 + hy, zc=zc, hx=hx, hy=hy, hz=max(z2 - z1, 8), f=0.0))
    elif mode == "HEX":
        Rhex = max(3.0, min(0.22 * min(float(nx_sc), float(ny_sc)), 0.5 * nx_sc - mg - 2.0))
        D.append(Defect("prism", xc=xc, yc=yc, zc=zc, R=Rhex, n=6, phi=_phi_cat(cfg, 0.0), f=0.0))
    elif mode == "HEX_STAG":
        Rhex = max(3.0, min(0.20 * min(float(nx_sc), float(ny_sc)), 0.5 * nx_sc - mg - 2.0))
        D.append(Defect("prism", xc=xc, yc=yc, zc=zc, R=Rhex, n=6, phi=_phi_cat(cfg, 0.0), f=0.0))
        Rtri = 4.0
        phi_t = _phi_cat(cfg, 0.0)
        xs = [oround(v) for v in _linspace(xsafe_lo + 4, xsafe_hi - 4, 4)]
        dyr = oround(0.22 * ny_sc)
        for s_, xv in enumerate(xs, start=1):
            ys = [yc - dyr, yc + dyr] if s_ % 2 == 1 else [yc - 2 * dyr, yc, yc + 2 * dyr]
            for yv in ys:
                if not (y1 + 2 < yv < y2 - 2):
                    continue
                if not math.hypot(float(xv - xc), float(yv - yc)) > Rhex + Rtri + 1:
                    continue
                D.append(Defect("prism", xc=xv, yc=oround(yv), zc=zc, R=Rtri, n=3, phi=phi_t, f=0.0))
    else:
        raise ValueError(f"unknown GEOM_MODE: {cfg.GEOM_MODE} (catalogue: {' | '.join(CATALOGUE)}, any of them "
                         f"with the suffix _REF; CUSTOM | CUSTOM_REF with geom_file)")
    if cfg.x_shift != 0:
        shift_defects(D, "x", oround(cfg.x_shift * nx_sc))
    if cfg.y_shift != 0:
        shift_defects(D, "y", oround(cfg.y_shift * ny_sc))
    if is_ref:
        mirror_defects(D, cfg, x1, x2, y1, y2)
    apply_zmode(D, cfg, z1, z2)
    return D


def margin_warnings(defects, cfg: GeomSpec, x1, x2) -> List[str]:
    """check_injection_margin (geometry.jl:580-596): the defects that reach closer than inj_margin
    cells to the injection faces (a warning: the current enters through the self-field)."""
    out = []
    if cfg.current_axis.lower() != "x":
        return out
    mg = cfg.inj_margin
    for d in defects:
        if d.type in ("sphere", "cylinder", "prism"):
            lo, hi = d.xc - math.ceil(d.R), d.xc + math.ceil(d.R)
        elif d.type == "box":
            lo, hi = d.xc - d.hx, d.xc + d.hx
        elif d.type == "wedge":
            lo, hi = d.xc - 2, d.xc + d.half
        elif d.type == "fgrad":
            lo, hi = d.xa, d.xb
        else:
            continue
        if lo < x1 + mg or hi > x2 - mg:
            out.append(f"[margin] defect '{d.type}' spans x={lo}..{hi}; the safe zone is {x1 + mg}..{x2 - mg} "
                       f"(inj_margin={mg})")
    return out


# ---------------------------------------------------------------------------- shape files
_Z_WORDS = ("all", "bottom", "top", "buried")
_TYPES = ("circle", "rect", "triangle", "ngon", "polygon", "sphere", "pixels", "voxels")
_OLD_KEYS = {"tipo": "type", "sesgo": "skew", "puntos": "points", "filas": "rows", "paleta": "palette", "capas": "layers"}
_OLD_TYPES = {"circulo": "circle", "triangulo": "triangle", "ngono": "ngon", "poligono": "polygon", "esfera": "sphere",
              "pixeles": "pixels", "voxeles": "voxels"}
_OLD_Z = {"todo": "all", "abajo": "bottom", "arriba": "top", "enterrado": "buried"}


def _norm_z(z):
    return _OLD_Z.get(z, z) if isinstance(z, str) else z


def normalise_element(e: dict) -> dict:
    """_normalise (shapes.jl:66-78): the first version's (Spanish) names to the current ones."""
    o = {_OLD_KEYS.get(k, k): v for k, v in e.items()}
    if "type" in o:
        o["type"] = _OLD_TYPES.get(o["type"], o["type"])
    if "z" in o:
        o["z"] = _norm_z(o["z"])
    if "palette" in o:
        o["palette"] = {ch: {k: (_norm_z(v) if k == "z" else v) for k, v in q.items()} for ch, q in o["palette"].items()}
    return o


def shape_elements(doc: dict, where: str = "shape file") -> List[dict]:
    """The element list of a parsed shape file (read_shape, shapes.jl:84-96)."""
    raw = doc.get("shape", doc.get("forma"))
    if raw is None:
        raise ValueError(f"{where}: no [[shape]] block")
    elems = [normalise_element(e) for e in raw]
    for n, e in enumerate(elems, start=1):
        if "type" not in e:
            raise ValueError(f"{where}: shape #{n} has no 'type'")
        if e["type"] not in _TYPES:
            raise ValueError(f"{where}: shape #{n} has unknown type '{e['type']}'")
    return elems


def read_shape(path: str) -> List[dict]:
    """read_shape (shapes.jl:84-96): a shape file (TOML) as its list of elements."""
    import tomli
    if not os.path.isfile(path):
        raise ValueError(f"geom_file does not exist: {path}")
    with open(path, "rb") as fh:
        return shape_elements(tomli.load(fh), path)


def _range_z(spec, Lz):
    """_rango_z (shapes.jl:100-112): the extent of a z (or `along`) specification, in ξσ."""
    if spec is None:
        return (-math.inf, math.inf)
    if isinstance(spec, str):
        if spec == "all":
            return (-math.inf, math.inf)
        if spec == "bottom":
            return (-math.inf, Lz / 2)
        if spec == "top":
            return (Lz / 2, math.inf)
        if spec == "buried":
            return (Lz / 4, 3 * Lz / 4)
        raise ValueError(f"z = '{spec}' is not one of {_Z_WORDS} nor [z0, z1]")
    if len(spec) != 2:
        raise ValueError(f"z = {spec}: expected [z0, z1]")
    return (float(spec[0]), float(spec[1]))


def _local(e, X, Y):
    """_local (shapes.jl:118-126): the point in the element's frame (centre at 0, rotation undone)."""
    cx, cy = float(e.get("x", 0.0)), float(e.get("y", 0.0))
    th = math.radians(float(e.get("rot", 0.0)))
    dx, dy = X - cx, Y - cy
    if th == 0.0:
        return dx, dy
    c, s = math.cos(th), math.sin(th)
    return c * dx + s * dy, -s * dx + c * dy


def in_polygon(px, py, vx, vy):
    """_en_poligono (shapes.jl:128-148), element-wise: the crossing number, a point on an edge
    (within the reference's 1e-9) inside."""
    n = len(vx)
    inside = np.zeros(np.shape(px), dtype=bool)
    on = np.zeros(np.shape(px), dtype=bool)
    j = n - 1
    for i in range(n):
        xi, yi, xj, yj = vx[i], vy[i], vx[j], vy[j]
        cr = (xj - xi) * (py - yi) - (yj - yi) * (px - xi)
        on |= ((np.abs(cr) <= TOL * max(1.0, math.hypot(xj - xi, yj - yi)))
               & (min(xi, xj) - TOL <= px) & (px <= max(xi, xj) + TOL)
               & (min(yi, yj) - TOL <= py) & (py <= max(yi, yj) + TOL))
        if yj != yi:  # a horizontal edge is crossed by no ray
            crosses = (yi > py) != (yj > py)
            xint = xi + (py - yi) * (xj - xi) / (yj - yi)
            inside ^= crosses & (px < xint)
        j = i
    return on | inside


def inside_element(e, X, Y, Z, L):
    """_dentro (shapes.jl:142-186), element-wise over the points (X, Y, Z)."""
    t = e["type"]
    if t == "sphere":
        cx, cy, cz, r = float(e["x"]), float(e["y"]), float(e["zc"]), float(e["r"])
        return (X - cx) ** 2 + (Y - cy) ** 2 + (Z - cz) ** 2 <= r * r + TOL
    ax = e.get("axis", "z")
    if ax not in ("x", "y", "z"):
        raise ValueError(f"axis = '{ax}' is not one of ('x', 'y', 'z')")
    if ax != "z":
        A, la, U, V = (X, L[0], Y, Z) if ax == "x" else (Y, L[1], X, Z)
        a0, a1 = _range_z(e.get("along", "all"), la)
        ok = (a0 - TOL <= A) & (A <= a1 + TOL)
        q = {k: v for k, v in e.items() if k not in ("x", "y", "z", "u", "v", "axis", "along")}
        q["x"] = e.get("u", 0.0)
        q["y"] = e.get("v", 0.0)
        return ok & inside_element(q, U, V, np.zeros_like(U), L)
    z0, z1 = _range_z(e.get("z", "all"), L[2])
    ok = (z0 - TOL <= Z) & (Z <= z1 + TOL)
    if t == "circle":
        cx, cy, r = float(e["x"]), float(e["y"]), float(e["r"])
        return ok & ((X - cx) ** 2 + (Y - cy) ** 2 <= r * r + TOL)
    if t == "rect":
        u, v = _local(e, X, Y)
        return ok & (np.abs(u) <= float(e["w"]) / 2 + TOL) & (np.abs(v) <= float(e["h"]) / 2 + TOL)
    if t == "triangle":
        u, v = _local(e, X, Y)
        w, h, s = float(e["w"]), float(e["h"]), float(e.get("skew", 0.0))
        return ok & in_polygon(u, v, (-w / 2, w / 2, s), (0.0, 0.0, h))
    if t == "ngon":
        u, v = _local(e, X, Y)
        n, r = int(e["n"]), float(e["r"])
        vx = tuple(r * math.cos(2 * math.pi * k / n) for k in range(n))
        vy = tuple(r * math.sin(2 * math.pi * k / n) for k in range(n))
        return ok & in_polygon(u, v, vx, vy)
    if t == "polygon":
        pts = e["points"]
        if len(pts) < 3:
            raise ValueError("polygon with fewer than 3 points")
        u, v = _local(e, X, Y)
        return ok & in_polygon(u, v, tuple(float(p[0]) for p in pts), tuple(float(p[-1]) for p in pts))
    raise ValueError(f"type '{t}' not handled here")


def _rows(t):
    if isinstance(t, str):
        if t.endswith("\n"):
            t = t[:-1]
        return t.split("\n")
    return list(t)


def _lookup(rows, palette, r, c, what):
    """The palette entry at 1-based row r, column c of `rows` (None where it leaves the node)."""
    nr = len(rows)
    if not (1 <= r <= nr):
        return None
    fila = rows[r - 1]
    if not (1 <= c <= len(fila)):
        return None
    ch = fila[c - 1]
    if ch in (".", " "):
        return None
    if ch not in palette:
        raise ValueError(f"{what}: character '{ch}' is not in the palette")
    return ch


def _value(q) -> float:
    """_valor (shapes.jl:211): the f an element or palette entry writes, NaN for vacuum."""
    if q.get("vacuum", False) is True:
        return float("nan")
    return float(q.get("f", 1.0 if "material" in q else 0.0))


def rasterize_shape(fmask, elems, gr, mirror=False, vac=None, mat=None, matnames=()):
    """rasterize_shape! (shapes.jl:245-287): each element's value on the nodes x1:x2, y1:y2,
    z1:z2 it covers, in order; the nodes at their physical position from the corner (x1, y1, z1),
    mirrored y -> (y1 + y2 + 1) - y for CUSTOM_REF.  Vacuum marks `vac` (f = 0), a later value gives
    the material back; a named material writes its index in `mat` (2 for the first name)."""
    x1, x2, y1, y2, z1, z2 = gr.x1, gr.x2, gr.y1, gr.y2, gr.z1, gr.z2
    dx, dy, dz = gr.dx, gr.dy, gr.dz
    L = ((x2 - x1) * dx, (y2 - y1) * dy, (z2 - z1) * dz)
    I, J, K = _block(x1, x2, y1, y2, z1, z2)
    X = (I - x1) * dx
    JJ = (y1 + y2 + 1) - J if mirror else J
    Y = (JJ - y1) * dy
    Z = (K - z1) * dz
    sub = (slice(x1 - 1, x2), slice(y1 - 1, y2), slice(z1 - 1, z2))
    F = fmask[sub]
    VAC = vac[sub] if vac is not None else None
    MAT = mat[sub] if mat is not None else None

    def write(m, q):
        if not m.any():
            return
        v = _value(q)
        if "material" in q:
            if MAT is None:
                raise ValueError(f"material = \"{q['material']}\": the run file has no [materials] tables")
            name = str(q["material"])
            if name not in matnames:
                raise ValueError(f"material = \"{name}\" is not a [materials.<name>] table of the run file "
                                 f"(there are: {', '.join(matnames)})")
            MAT[m] = matnames.index(name) + 2
        if math.isnan(v):
            if VAC is None:
                raise ValueError("vacuum = true needs release 2 and an insulating padding (sigma_vac)")
            VAC[m] = True
            F[m] = 0.0
        else:
            if VAC is not None:
                VAC[m] = False
            F[m] = v

    for e in elems:
        if e["type"] in ("pixels", "voxels"):
            if e["type"] == "pixels":
                d = float(e.get("dx", 0.5))
                rows_of = {0: _rows(e["rows"])}
                lidx = np.zeros(X.shape, dtype=np.int64)
                dzv = None
            else:
                d = float(e.get("dx", 0.5))
                dzv = float(e.get("dz", d))
                rows_of = {l: _rows(t) for l, t in enumerate(e["layers"])}
                lidx = np.rint(Z / dzv).astype(np.int64)  # round(Int, Z / dz) (ties to even), 0-based layer
            cidx = np.rint(X / d).astype(np.int64) + 1
            ridx_off = np.rint(Y / d).astype(np.int64)
            pal = e["palette"]
            chars = {}
            for pos in np.ndindex(X.shape):
                l = int(lidx[pos])
                if l not in rows_of:
                    continue
                rows = rows_of[l]
                ch = _lookup(rows, pal, len(rows) - int(ridx_off[pos]), int(cidx[pos]), e["type"])
                if ch is not None:
                    chars.setdefault(ch, []).append(pos)
            for ch, poss in chars.items():
                q = pal[ch]
                m = np.zeros(X.shape, dtype=bool)
                idx = tuple(np.array(poss).T)
                m[idx] = True
                if e["type"] == "pixels":
                    z0, z1r = _range_z(q.get("z", "all"), L[2])
                    m &= (z0 - TOL <= Z) & (Z <= z1r + TOL)
                # the palette entries of one layer write in node order, each node once: their order
                # among themselves does not matter
                write(m, q)
            continue
        m = inside_element(e, X, Y, Z, L)
        if e.get("outside", False) is True:
            m = ~m
        write(m, e)
    return fmask


def ghost_planes(v, gr):
    """_ghost_planes! (geometry.jl:529-535): the six face ghost planes copy the plane inside."""
    x1, x2, y1, y2, z1, z2 = gr.x1, gr.x2, gr.y1, gr.y2, gr.z1, gr.z2
    v[x1 - 1] = v[x1]
    v[x2] = v[x2 - 1]
    v[:, y1 - 1] = v[:, y1]
    v[:, y2] = v[:, y2 - 1]
    v[:, :, z1 - 1] = v[:, :, z1]
    v[:, :, z2] = v[:, :, z2 - 1]
    return v


def sculpt(mk, gr):
    """sculpt! (geometry.jl:537-551): the vacuum nodes (with their face ghosts) leave the block:
    τψ = f = 0 on them and every link that touches one is cut."""
    v = ghost_planes(mk.vac, gr)
    mk.tpsi[v] = 0.0
    mk.f[v] = 0.0
    mk.tx[v] = 0.0
    mk.ty[v] = 0.0
    mk.tz[v] = 0.0
    mk.tx[:-1][v[1:]] = 0.0
    mk.ty[:, :-1][v[:, 1:]] = 0.0
    mk.tz[:, :, :-1][v[:, :, 1:]] = 0.0
    x1, x2, y1, y2, z1, z2 = gr.x1, gr.x2, gr.y1, gr.y2, gr.z1, gr.z2
    if not (mk.tpsi[x1:x2, y1:y2, z1:z2] > 0).any():
        raise ValueError("the shape leaves no material: everything is vacuum")
    return mk


def conducting_nodes(gr, vac=None):
    """conducting_nodes (geometry.jl:559-570): the bar y1:y2+1 x z1:z2+1 along the whole x, minus
    the vacuum of a sculpted sample (the contacts take the section of the face they touch)."""
    Nx, Ny, Nz = gr.Nx, gr.Ny, gr.Nz
    c = np.zeros((Nx + 1, Ny + 1, Nz + 1), dtype=bool)
    c[:, gr.y1 - 1:gr.y2 + 1, gr.z1 - 1:gr.z2 + 1] = True
    if vac is None:
        return c
    ii = np.clip(np.arange(1, Nx + 2), gr.x1, gr.x2 + 1) - 1
    sec = vac[ii][:, gr.y1 - 1:gr.y2 + 1, gr.z1 - 1:gr.z2 + 1]
    c[:, gr.y1 - 1:gr.y2 + 1, gr.z1 - 1:gr.z2 + 1] &= ~sec
    return c


def shape_margin(fmask, cfg: GeomSpec, gr) -> int:
    """margen_forma (shapes.jl:329-346): the nodes with f < 1 closer than inj_margin cells to the
    injection faces."""
    x1, x2, y1, y2, z1, z2 = gr.x1, gr.x2, gr.y1, gr.y2, gr.z1, gr.z2
    mg = cfg.inj_margin
    sub = fmask[x1 - 1:x2, y1 - 1:y2, z1 - 1:z2] < 1
    if cfg.current_axis.lower() == "x":
        i = np.arange(x1, x2 + 1)
        near = (i < x1 + mg) | (i > x2 - mg)
        return int(sub[near].sum())
    j = np.arange(y1, y2 + 1)
    near = (j < y1 + mg) | (j > y2 - mg)
    return int(sub[:, near].sum())


@dataclass
class Shape:
    """What build_masks returns beyond the box masks: the vacuum of a sculpted sample and the
    material index per node (None where the shape has none), with the margin warnings."""
    vac: Optional[np.ndarray] = None
    mat: Optional[np.ndarray] = None
    warnings: List[str] = field(default_factory=list)


def build_masks(cfg: GeomSpec, gr, elems: Optional[List[dict]] = None):
    """build_masks (geometry.jl:485-527): the box masks with the catalogue's defects painted, or
    the shape file's elements rasterised (CUSTOM, CUSTOM_REF: `elems`, else read from
    cfg.geom_file).  Returns (Masks, Shape)."""
    from .scheme import box_masks
    mk = box_masks(gr)
    mode = cfg.GEOM_MODE.upper()
    x1, x2, y1, y2, z1, z2 = gr.x1, gr.x2, gr.y1, gr.y2, gr.z1, gr.z2
    if mode in ("CUSTOM", "CUSTOM_REF"):
        if elems is None:
            if not cfg.geom_file:
                raise ValueError(f"GEOM_MODE={mode} needs geom_file")
            elems = read_shape(cfg.geom_file)
        sh = (gr.Nx + 1, gr.Ny + 1, gr.Nz + 1)
        vac = np.zeros(sh, dtype=bool, order="F")
        has_mat = any("material" in e or ("palette" in e and any("material" in q for q in e["palette"].values()))
                      for e in elems)
        mat = np.ones(sh, dtype=np.uint8, order="F") if has_mat else None
        rasterize_shape(mk.f, elems, gr, mirror=mode == "CUSTOM_REF", vac=vac, mat=mat, matnames=list(cfg.materials))
        if mat is not None:
            ghost_planes(mat, gr)
            mk.mat = mat
        out = Shape(mat=mat)
        if vac.any():
            mk.vac = vac
            sculpt(mk, gr)
            out.vac = mk.vac
            return mk, out
        n = shape_margin(mk.f, cfg, gr)
        if n:
            out.warnings.append(f"[margin] the shape touches {n} node(s) closer than {cfg.inj_margin} cells to the "
                                f"injection faces")
        return mk, out
    defects = build_defect_list(cfg, x1, x2, y1, y2, z1, z2)
    for d in defects:
        paint_defect(mk.f, d, x1, x2, y1, y2, z1, z2)
    return mk, Shape(warnings=margin_warnings(defects, cfg, x1, x2))
