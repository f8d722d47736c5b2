"""Checkpoints of a run and device snapshots.

A checkpoint is the reference's (gtdgl/src/io.jl:1-8, 131-150): every band, the link phases of this
step and of the previous one, and the field, under the reference's keys

    tdgl_version, nbands, G1..GN, thetax/y/z, thetax_prev/y/z (φ after advance!'s swap), hx/y/z

plus the `extra` entries a ladder stores with it: `open_extra` (io.jl:159-161), the warm PCG
history (`pcg_x1..3`, `pcg_p1..3`, protocol.jl:656-658) and anything else the caller passes
(`jp`, `J`, ..., protocol.jl:627-660).  The container is a NumPy `.npz` (compressed, one `.npy` per
key, arrays in the reference's column-major shapes): HDF5/JLD2 has no writer in this environment.
`read_state` also opens the reference's own checkpoints, JLD2 (juliatdgl_amd.jld2) and the Octave
text `.mat` of the releases up to 1.1 (io.jl:19-97), so a run of the reference resumes here.

Resuming from a checkpoint rebuilds the links V = e^{-iθ}, U = e^{-iθ_prev} on the host, as the
reference's `load_state` (io.jl:181-191, state.jl:69-79); a device snapshot (`snapshot` /
`restore`) keeps every device array instead, so that continuing from it is continuing to the bit.
"""
from __future__ import annotations

import os
import re
from dataclasses import dataclass
from typing import Optional

import numpy as np

from .scheme import Grid
from .state import State

CODE_VERSION = "TDGL 2.0 (juliatdgl_amd)"


# ----------------------------------------------------------------------------- writing
def _entry(v):
    """A value as an array np.savez stores without pickling."""
    if isinstance(v, np.ndarray):
        return v
    if isinstance(v, str):
        return np.array(v)
    if isinstance(v, (bool, np.bool_)):
        return np.array(bool(v))
    if isinstance(v, (int, np.integer)):
        return np.array(int(v), dtype=np.int64)
    if isinstance(v, (float, np.floating)):
        return np.array(float(v))
    if isinstance(v, (list, tuple)):
        return np.asarray(v)
    raise TypeError(f"save_state: {type(v).__name__} is not storable")


def state_pairs(s: State, prefix: str = "") -> dict:
    """The state's entries under the reference's names (_state_pairs, protocol.jl:618-621)."""
    out = {f"{prefix}G{mu + 1}": s.G[mu] for mu in range(s.N)}
    out.update({f"{prefix}thetax": s.thx, f"{prefix}thetay": s.thy, f"{prefix}thetaz": s.thz,
                f"{prefix}thetax_prev": s.phx, f"{prefix}thetay_prev": s.phy, f"{prefix}thetaz_prev": s.phz,
                f"{prefix}hx": s.hx, f"{prefix}hy": s.hy, f"{prefix}hz": s.hz})
    return out


def save_state(path: str, s: State, **extra) -> str:
    """save_state(path, s; extra...) (io.jl:131-150): written atomically (a temporary file, then
    a rename)."""
    d = os.path.dirname(os.path.abspath(path))
    os.makedirs(d, exist_ok=True)
    ent = {"tdgl_version": CODE_VERSION, "nbands": s.N}
    ent.update(state_pairs(s))
    ent.update(extra)
    tmp = path + ".tmp"
    with open(tmp, "wb") as fh:
        np.savez_compressed(fh, **{k: _entry(v) for k, v in ent.items()})
    os.replace(tmp, path)
    return path


def checkpoint_path(d: str, label: str) -> str:
    """checkpoint_path(dir, label) (io.jl:117-121): state_<label> with the extension that exists
    (the newer if several do), else `.npz`."""
    cands = [os.path.join(d, f"state_{label}{e}") for e in (".npz", ".jld2", ".mat")]
    have = [p for p in cands if os.path.isfile(p)]
    if not have:
        return cands[0]
    return max(have, key=os.path.getmtime)


# ----------------------------------------------------------------------------- reading
def _scalar(v):
    if isinstance(v, np.ndarray) and v.ndim == 0:
        x = v.item()
        return x
    return v


def read_npz(path: str) -> dict:
    with np.load(path, allow_pickle=False) as z:
        return {k: _scalar(z[k]) for k in z.files}


def read_jld2(path: str) -> dict:
    """A JLD2 checkpoint of the reference (read_jld2, io.jl:110)."""
    from .jld2 import JLD2File
    f = JLD2File(path)
    return {k: f[k] for k in f.keys() if k != "_types"}


def _complex(t: str) -> complex:
    """An Octave complex `(re,im)` (_parse_complejo, io.jl:88-94)."""
    t = t.strip()
    if not (t.startswith("(") and t.endswith(")")):
        raise ValueError(f"octmat: '{t}' is not an Octave complex number")
    re_, im_ = t[1:-1].split(",")
    return complex(float(re_), float(im_))


def read_octave_text(path: str) -> dict:
    """The text format of Octave's `save` (read_octave_text, io.jl:19-73): scalar, bool, complex
    scalar, string, and (bool / complex) matrices with `# rows/# columns` or `# ndims`, data in
    column order."""
    with open(path, "r") as fh:
        lines = fh.read().splitlines()
    out = {}
    i, n = 0, len(lines)
    while i < n:
        line = lines[i]
        if not line.startswith("# name: "):
            i += 1
            continue
        name = line[8:].strip()
        i += 1
        if i >= n or not lines[i].startswith("# type: "):
            raise ValueError(f"octmat: '{name}' without '# type:'")
        typ = lines[i][8:].strip()
        i += 1
        if typ == "scalar":
            out[name] = float(lines[i].strip())
            i += 1
        elif typ == "bool":
            out[name] = float(lines[i].strip()) != 0
            i += 1
        elif typ == "complex scalar":
            out[name] = _complex(lines[i])
            i += 1
        elif typ == "string":
            i += 1  # '# elements:'
            ln = int(lines[i].split(":")[1])
            i += 1
            out[name] = lines[i][:ln]
            i += 1
        elif typ in ("matrix", "complex matrix", "bool matrix"):
            if lines[i].startswith("# ndims:"):
                dims = [int(x) for x in lines[i + 1].split()]
                i += 2
            elif lines[i].startswith("# rows:"):
                r = int(lines[i].split(":")[1])
                if not lines[i + 1].startswith("# columns:"):
                    raise ValueError(f"octmat: '{name}' has '# rows:' but no '# columns:'")
                dims = [r, int(lines[i + 1].split(":")[1])]
                i += 2
            else:
                raise ValueError(f"octmat: '{name}' without dimensions")
            tot = int(np.prod(dims))
            cplx = typ.startswith("complex")
            vals = []
            while len(vals) < tot:
                t = lines[i].strip()
                i += 1
                if not t:
                    continue
                if cplx:
                    vals.append(_complex(t))
                else:
                    vals.extend(float(x) for x in t.split())
            arr = np.array(vals[:tot], dtype=np.complex128 if cplx else np.float64)
            out[name] = arr.reshape(dims, order="F")
        else:
            raise ValueError(f"octmat: unsupported type '{typ}' in '{name}'")
    return out


def read_state(path: str) -> dict:
    """read_state(path) (io.jl:100-108): every entry of a checkpoint, by its extension; a binary
    MATLAB file is refused as the reference refuses it."""
    if not os.path.isfile(path):
        raise FileNotFoundError(f"does not exist: {path}")
    low = path.lower()
    if low.endswith(".npz"):
        return read_npz(path)
    if low.endswith(".jld2"):
        return read_jld2(path)
    with open(path, "rb") as fh:
        if fh.read(6) == b"MATLAB":
            raise ValueError(f"{path} is a binary MATLAB file: read it with release 1.2 and save it as JLD2")
    return read_octave_text(path)


def load_state(path: str, gr: Grid):
    """load_state(path, gr) -> (State, entries) (io.jl:181-191): without the previous link phases
    (Octave text, old JLD2) they are taken equal to the current ones."""
    w = read_state(path)
    N = int(w["nbands"]) if "nbands" in w else 2
    sz = (gr.Nx + 1, gr.Ny + 1, gr.Nz + 1)
    G = [np.asarray(w[f"G{mu + 1}"], dtype=np.complex128) for mu in range(N)]
    for g in G:
        if g.shape != sz:
            raise ValueError(f"{path}: Ψ is {g.shape}, the grid needs {sz}")
    f64 = lambda k: np.asarray(w[k], dtype=np.float64)
    th = tuple(f64(f"theta{c}") for c in "xyz")
    thp = tuple(f64(f"theta{c}_prev") for c in "xyz") if "thetax_prev" in w else th
    h = tuple(f64(f"h{c}") for c in "xyz")
    return State.from_fields(gr, G, th, thp, h), w


# ----------------------------------------------------------------------------- stepper-side extras
def open_extra(dev) -> dict:
    """open_extra(st) (io.jl:159-161) from the device: the clock and the ramp of the open boundary,
    empty for the other boundaries."""
    if dev.st.selffield_bc != "open":
        return {}
    count, old, new = dev.get_open()
    out = {"open_count": int(count)}
    for c, x in enumerate("xyz"):
        out[f"open_old_{x}"] = old[c]
        out[f"open_new_{x}"] = new[c]
    return out


def restore_open(dev, w: dict) -> None:
    """restore_open!(st, w) (io.jl:163-171)."""
    if dev.st.selffield_bc != "open" or "open_count" not in w:
        return
    dev.set_open(int(w["open_count"]), [w[f"open_old_{x}"] for x in "xyz"], [w[f"open_new_{x}"] for x in "xyz"])


def pcg_extra(dev) -> dict:
    """The warm PCG history a ladder checkpoint holds (protocol.jl:656-658): the last solution,
    and with it the one before (the extrapolation of the initial guess), when there are."""
    if dev.st.field_mode != 2:
        return {}
    p = dev.get_pcg()
    out = {}
    if p.warm:
        out.update({f"pcg_x{c + 1}": p.x[c] for c in range(3)})
        if p.have2:
            out.update({f"pcg_p{c + 1}": p.xprev[c] for c in range(3)})
    return out


def restore_pcg(dev, w: dict) -> None:
    """_pcg_back! of a resumed ladder (protocol.jl:380-381, implicit.jl:198-207)."""
    if dev.st.field_mode != 2 or "pcg_x1" not in w:
        return
    p = dev.get_pcg()
    p.warm = True
    for c in range(3):
        p.x[c][...] = np.asarray(w[f"pcg_x{c + 1}"], dtype=np.float64)
    p.have2 = "pcg_p1" in w
    if p.have2:
        for c in range(3):
            p.xprev[c][...] = np.asarray(w[f"pcg_p{c + 1}"], dtype=np.float64)
    dev.set_pcg(p)


def save_checkpoint(path: str, dev, **extra) -> str:
    """The device's state with the extras resuming needs (the open ramp, the PCG history, the
    noise stream when the stepper has noise) and the caller's: one file, written at once."""
    s = dev.download()
    ent = dict(open_extra(dev))
    ent.update(pcg_extra(dev))
    if hasattr(dev, "noise_extra"):
        ent.update(dev.noise_extra())
    ent.update(extra)
    return save_state(path, s, **ent)


def resume_checkpoint(path: str, dev):
    """Load a checkpoint into a fresh device stepper: the state (links rebuilt from θ), then
    what save_checkpoint stored with it.  Returns the entries of the file."""
    s, w = load_state(path, dev.st.grid)
    if s.N != dev.st.N:
        raise ValueError(f"{path}: {s.N} bands, the stepper has {dev.st.N}")
    dev.upload(s)
    restore_open(dev, w)
    restore_pcg(dev, w)
    if hasattr(dev, "noise_restore"):
        dev.noise_restore(w)
    return w


# ----------------------------------------------------------------------------- device snapshots
@dataclass
class Snapshot:
    """Everything the device holds of a run between two steps: every array of the state (links
    and F included), the PCG history, the open boundary's clock and ramp, the noise stream."""
    state: State
    pcg: Optional[object]
    open: Optional[tuple]
    noise: Optional[dict]


def snapshot(dev) -> Snapshot:
    st = dev.st
    return Snapshot(dev.download(), dev.get_pcg() if st.field_mode == 2 else None,
                    dev.get_open() if st.selffield_bc == "open" else None,
                    dev.noise_extra() if hasattr(dev, "noise_extra") else None)


def restore(dev, sn: Snapshot) -> None:
    """Continue from a snapshot on this stepper (a fresh one of the same Stepper, or the same one
    after other work): to the bit."""
    dev.upload(sn.state)
    if sn.pcg is not None:
        dev.set_pcg(sn.pcg)
    if sn.open is not None:
        dev.set_open(*sn.open)
    if sn.noise is not None and hasattr(dev, "noise_restore"):
        dev.noise_restore(sn.noise)


def save_snapshot(path: str, sn: Snapshot) -> str:
    """A snapshot as a file: the checkpoint keys, and every further device array under its
    state field name (F1.., Vx.., Ux.., Jx..) so that `load_snapshot` continues to the bit."""
    s = sn.state
    ent = {f"F{mu + 1}": s.F[mu] for mu in range(s.N)}
    for k in ("Vx", "Vy", "Vz", "Ux", "Uy", "Uz", "Jx", "Jy", "Jz"):
        ent[k] = getattr(s, k)
    if sn.pcg is not None:
        p = sn.pcg
        ent.update({f"snap_pcg_x{c + 1}": p.x[c] for c in range(3)})
        ent.update({f"snap_pcg_p{c + 1}": p.xprev[c] for c in range(3)})
        ent.update(snap_pcg_warm=bool(p.warm), snap_pcg_have2=bool(p.have2), snap_pcg_iters=int(p.iters))
    if sn.open is not None:
        count, old, new = sn.open
        ent["open_count"] = int(count)
        for c, x in enumerate("xyz"):
            ent[f"open_old_{x}"] = old[c]
            ent[f"open_new_{x}"] = new[c]
    if sn.noise is not None:
        ent.update(sn.noise)
    return save_state(path, s, **ent)


def load_snapshot(path: str, dev) -> Snapshot:
    s, w = load_state(path, dev.st.grid)
    for mu in range(s.N):
        s.F[mu] = np.asfortranarray(w[f"F{mu + 1}"])
    for k in ("Vx", "Vy", "Vz", "Ux", "Uy", "Uz", "Jx", "Jy", "Jz"):
        setattr(s, k, np.asfortranarray(w[k]))
    pcg = None
    if "snap_pcg_x1" in w:
        pcg = dev.get_pcg()
        for c in range(3):
            pcg.x[c][...] = w[f"snap_pcg_x{c + 1}"]
            pcg.xprev[c][...] = w[f"snap_pcg_p{c + 1}"]
        pcg.warm, pcg.have2, pcg.iters = bool(w["snap_pcg_warm"]), bool(w["snap_pcg_have2"]), int(w["snap_pcg_iters"])
    op = None
    if "open_count" in w:
        op = (int(w["open_count"]), [np.asfortranarray(w[f"open_old_{x}"]) for x in "xyz"],
              [np.asfortranarray(w[f"open_new_{x}"]) for x in "xyz"])
    noise = {k: v for k, v in w.items() if re.match(r"noise_", k)} or None
    return Snapshot(s, pcg, op, noise)
