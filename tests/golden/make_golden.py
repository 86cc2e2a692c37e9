"""Writes tests/golden/tiny_plain.npz: the "tiny" seeded case (6 holograms
64^2 float32, crop 16^2, 32^3 volume, no mask, C_ABER 0: a 16^2 crop leaves
16 fit samples for the 15 unknowns of the aberration fit, too few for a
well-conditioned target) with the FP64 oracle's speculars, fields and index
volume. The reference ships no golden vectors;
these fix the oracle's outputs so that a change to it shows, and give the
GPU path a stored target. Run from the repository root:
    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, ".."))
sys.path.insert(0, os.path.join(HERE, "..", ".."))

from cases import make_case  # noqa: E402
from oracle import ref  # noqa: E402


def main():
    case = make_case("tiny", c_aber=0)
    holos = case.acq.holos
    fields, pos = ref.preprocess(holos, case.g, None, "f64")
    index, _, cnt, _ = ref.reconstruct(fields, pos, case.g, "f64")
    p = case.params
    np.savez_compressed(
        os.path.join(HERE, "tiny_plain.npz"),
        holos=holos.astype(np.float32), posspec=pos.astype(np.int32),
        fields=fields.astype(np.complex128), index=index.astype(np.complex64),
        counts_total=np.int64(cnt.sum()),
        geom=np.array([p["n"], p["nxmax"], p["v"], p["cx"], p["cy"], p["c_aber"]],
                      dtype=np.int64),
        pix=np.float64(p["pix"]))


if __name__ == "__main__":
    main()
