"""Convert the reference's golden JLD2 states into ``.npz`` fixtures.

    python tests/golden/convert_goldens.py [/root/reference]

The reference (gtdgl/test/golden/make_golden.jl) advanced a juliaTDGL 1.2 run N0 = 200 steps,
saved the whole state ("ini"), took N1 = 40 more steps and saved it again ("fin"), together with
the masks and the parameters of the step 1.2 chose.  gtdgl/test/test_regression12.jl:59-80 loads
"ini", takes N1 release-2.0 steps and requires "fin" to 1e-10 of max(1, |x|).

Each fixture keeps exactly the data that test reads: the parameters (as JSON), the masks, the
initial fields (Ψ, θ, φ, h, and the links the test checks) and the final fields (Ψ, θ, φ, V, h).
Arrays keep their Julia (column-major) shapes.  This is data, not code: the fixtures hold the
reference's numbers, the reader is juliatdgl_amd/jld2.py (written from the HDF5 specification).
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", ".."))
from juliatdgl_amd.jld2 import JLD2File  # noqa: E402

CASES = ("k2_R13_explicit", "k2_R4_prod", "k25_R4_prod")
INI = ("G1", "G2", "thetax", "thetay", "thetaz", "phix", "phiy", "phiz", "hx", "hy", "hz",
       "Vx", "Ux", "Uz")
FIN = ("G1", "G2", "thetax", "thetay", "thetaz", "phix", "phiy", "phiz", "Vx", "Vy", "Vz",
       "hx", "hy", "hz")
MASKS = ("fmask", "taux", "tauy", "tauz", "taupsi")


def convert(ref_root: str, out_dir: str) -> None:
    gdir = os.path.join(ref_root, "gtdgl", "test", "golden")
    for case in CASES:
        f = JLD2File(os.path.join(gdir, case + ".jld2"))
        par = f["par"]
        ini = f["ini"]
        fin = f["fin"]
        arrays = {}
        for k in INI:
            arrays["ini_" + k] = np.asfortranarray(ini[k])
        for k in FIN:
            arrays["fin_" + k] = np.asfortranarray(fin[k])
        for k in MASKS:
            arrays[k] = np.asfortranarray(f[k])
        par = {k: (v.item() if hasattr(v, "item") else v) for k, v in par.items()}
        arrays["par_json"] = np.frombuffer(json.dumps(par, sort_keys=True).encode(), dtype=np.uint8)
        out = os.path.join(out_dir, case + ".npz")
        np.savez_compressed(out, **arrays)
        print(f"{case}: dt={par['dt']} n_pi={par['n_pi']} -> {out} ({os.path.getsize(out)} B)")


if __name__ == "__main__":
    root = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    convert(root, os.path.dirname(os.path.abspath(__file__)))
