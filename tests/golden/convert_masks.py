"""Fixtures of the run-file front end from the reference's golden files.

    python tests/golden/convert_masks.py [/root/reference]

* catalogue_masks.npz: the condensation masks (fmask) that juliaTDGL 1.2 wrote for the catalogue
  modes the front end builds (gtdgl/test/golden/masks.jld2, made by make_masks.jl), with the mesh
  and geometry keys of each case as JSON.
* ladder_k25_scalars.json: the scalars of gtdgl/test/golden/ladder_k25.jld2 that
  gtdgl/test/test_config.jl compares a run file against (mr, dt, n_pi), with eps_c and the brackets.
* ladder_k25.npz: the two polarities of that golden ladder, point by point (J, verdicts, cycles, V,
  the band-2 slope, the static voltage, the half-window slopes, slips, magnetization, residuals)
  and the final Ψ of each, which gtdgl/test/test_protocol.jl:4-29 compares a ladder against.

Data only: the masks and numbers the reference's files hold, read with juliatdgl_amd/jld2.py.
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", ".."))
from juliatdgl_amd.jld2 import JLD2File  # noqa: E402

CASES = ("SIM_40x24x6", "SIM_REF_30x20x8", "WEDGE_40x24x6", "WEDGE_30x20x8", "WEDGE_REF_40x24x6",
         "WEDGE_REF_30x20x8", "WEDGE_BOTH_40x24x6", "WEDGE_BOTH_REF_30x20x8", "WEDGE_deep")


def main(ref: str) -> None:
    out = os.path.dirname(os.path.abspath(__file__))
    gdir = os.path.join(ref, "gtdgl", "test", "golden")
    c = JLD2File(os.path.join(gdir, "masks.jld2"))["casos"]
    arrs, meta = {}, {}
    for k in CASES:
        arrs[k] = np.asarray(c[k]["fmask"], dtype=np.float64)
        meta[k] = {"malla": c[k]["malla"], "geo": c[k]["geo"]}
    np.savez_compressed(os.path.join(out, "catalogue_masks.npz"), meta_json=np.frombuffer(json.dumps(meta).encode(), np.uint8),
                        **arrs)
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
