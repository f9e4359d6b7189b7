"""Writes tests/golden/vgrid_ref.npz: the icosahedral Voronoi velocity grid as
the reference's own icosahedron.h computes it, for nn = 1, 2, 3, 4, 8.

The values come from oracle/_ref/ref_vgrid, which `make -C oracle` compiles
from the reference's icosahedron.h and Point.h where they lie (the rest of the
reference cannot be built here: common.h requires <mpi.h>).  Run from the
repository root:  python tests/golden/make_vgrid_fixtures.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402

NNS = (1, 2, 3, 4, 8)


def main():
    out = {}
    for nn in NNS:
        g = O.ref_vgrid(nn)
        if g is None:
            raise SystemExit("oracle/_ref/ref_vgrid is not built (needs the reference sources)")
        for name, arr in zip(("vx", "vy", "vz", "vw"), g):
            out[f"{name}_{nn}"] = arr
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "vgrid_ref.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
