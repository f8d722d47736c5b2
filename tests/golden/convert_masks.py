"""Fixtures of the geometry and the run-file front end from the reference's golden files.

    python tests/golden/convert_masks.py [/root/reference]

* catalogue_masks.npz: every condensation mask (fmask) that juliaTDGL 1.2 wrote in
  gtdgl/test/golden/masks.jld2 (make_masks.jl: every catalogue mode and its _REF mirror on two
  meshes, the z modes, the shifts, the prism angle, the FGRAD keys, and every shape file of
  tdgl/shapes as CUSTOM and CUSTOM_REF), the box's link and node masks of its one case, and as
  JSON the mesh and geometry keys of each case and the parsed content of each shape file.
* ladder_k25_scalars.json, ladder_k25.npz: the golden ladder (gtdgl/test/golden/ladder_k25.jld2)
  that gtdgl/test/test_config.jl and test_protocol.jl compare against.

Data only: the masks, numbers and shape descriptions the reference's files hold, read with
juliatdgl_amd/jld2.py and tomli.
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", ".."))
from juliatdgl_amd.jld2 import JLD2File  # noqa: E402


def main(ref: str) -> None:
    import tomli
    out = os.path.dirname(os.path.abspath(__file__))
    gdir = os.path.join(ref, "gtdgl", "test", "golden")
    c = JLD2File(os.path.join(gdir, "masks.jld2"))["casos"]
    arrs, meta, shapes = {}, {}, {}
    for k in sorted(c.keys()):
        if k.startswith("_box"):
            for t in ("taux", "tauy", "tauz", "taupsi"):
                arrs[f"{k}_{t}"] = np.asarray(c[k][t], dtype=np.uint8)
            continue
        arrs[k] = np.asarray(c[k]["fmask"], dtype=np.float64)
        meta[k] = {"malla": c[k]["malla"], "geo": c[k]["geo"]}
        gf = c[k]["geo"].get("geom_file")
        if gf and gf not in shapes:
            with open(os.path.join(ref, "tdgl", "shapes", gf), "rb") as fh:
                shapes[gf] = tomli.load(fh)
    blob = json.dumps({"cases": meta, "shapes": shapes}).encode()
    np.savez_compressed(os.path.join(out, "catalogue_masks.npz"), meta_json=np.frombuffer(blob, np.uint8), **arrs)
    f = JLD2File(os.path.join(gdir, "ladder_k25.jld2"))
    sc = {k: f[k] for k in ("mr", "dt", "n_pi", "eps_c", "lo_pos", "hi_pos", "lo_neg", "hi_neg")}
    lad = {}
    for pol in ("pos", "neg"):
        d = f[pol]
        for k in ("J", "V", "V2", "Vst", "Vh", "Vh2", "nslip", "M", "r1", "r2", "m", "G1", "G2"):
            lad[f"{pol}_{k}"] = np.asarray(d[k])
        lad[f"{pol}_conv"] = np.frombuffer(d["conv"].encode(), np.uint8)
    np.savez_compressed(os.path.join(out, "ladder_k25.npz"), **lad)
    with open(os.path.join(out, "ladder_k25_scalars.json"), "w") as fh:
        json.dump({k: (float(v) if not isinstance(v, int) else v) for k, v in sc.items()}, fh)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
