"""Writes the golden fixtures of this directory.

tiny_iterations_dgz.npz: three source iterations of a tiny problem (4^3
zones in 2x1x2 zone sets, 2 group sets of 2 groups, GL 4x4 directions in 8
direction sets, P2 = 9 moments, DGZ, seeded boxes, seed 7) computed by the
numpy brute force of tests/bruteforce.py (einsum LTimes/LPlusTimes, dense
scattering definition, the sweep as one global linear solve of the
diamond-difference equations, math.fsum population).  The oracle only hosts
the generated inputs and the numpy state between iterations.

Stored: the config, per-iteration population, per-iteration phi of every PR
chunk in canonical [moment][group][zone] order, the final psi of every
subdomain in canonical [direction][group][zone] order, and sigt/volume of
zone set 0 (generator pin).

    python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "tests"), os.path.join(ROOT, "oracle"), ROOT]

import pyoracle as po  # noqa: E402
from bruteforce import numpy_iteration, population_fsum  # noqa: E402
from helpers import LAYOUT_AXES, canonical  # noqa: E402

CONFIG = dict(zones=(4, 4, 4), groups=4, quad=(4, 4), legendre=2, layout="DGZ", dset=8, gset=2,
              zset=(2, 1, 2), geometry=1, seed=7)
ITERATIONS = 3
M2L = np.array([0, 1, 1, 1, 2, 2, 2, 2, 2])


def main():
    layout = CONFIG["layout"]
    n = po.Oracle(po.make_input(**CONFIG), kind="ref", omp=False)
    n.init_psi()
    S0 = n.sdom(0)
    nz = S0["ni"] * S0["nj"] * S0["nk"]
    M = len(M2L)
    out = {"config": np.array(json.dumps(CONFIG)), "iterations": np.array(ITERATIONS)}
    pops = []
    for it in range(ITERATIONS):
        phi, psi = numpy_iteration(n, layout, M, M2L)
        for s, v in psi.items():
            n.set("psi", s, np.transpose(v, [("agz").index(ch) for ch in LAYOUT_AXES[layout]]).ravel())
        pops.append(population_fsum(n, layout))
        for c in range(n.num_pr):
            out["phi_it%d_pr%d" % (it, c)] = phi[c]
    for s in range(n.num_sdoms):
        out["psi_final_s%d" % s] = canonical(n.get("psi", s), layout, S0["nd"], S0["ng"], nz)
    out["population"] = np.array(pops)
    out["sigt_zonal_pr0"] = n.get("sigt_zonal", 0)
    out["volume_r0"] = n.get("volume", 0)
    path = os.path.join(HERE, "tiny_iterations_dgz.npz")
    np.savez_compressed(path, **out)
    print("wrote %s: populations %s" % (path, pops))


if __name__ == "__main__":
    main()
