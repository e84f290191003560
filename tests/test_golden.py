"""Committed golden fixtures (tests/golden/, written by make_golden.py from
the numpy brute force) against the CPU oracle in every nesting, and against
the HIP product through the C-ABI (gpu marker)."""
import json
import os

import numpy as np
import pytest

import pyoracle as po
from helpers import LAYOUT_AXES, canonical, rel_normwise

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = np.load(os.path.join(HERE, "golden", "tiny_iterations_dgz.npz"))
CONFIG = json.loads(str(FIX["config"]))
ITERS = int(FIX["iterations"])
M = 9
TOL = 1e-12


def _cfg(layout):
    c = dict(CONFIG)
    c["layout"] = layout
    c["zones"] = tuple(c["zones"])
    c["zset"] = tuple(c["zset"])
    c["quad"] = tuple(c["quad"])
    return c


def _check(get, num_pr, num_sdoms, layout, it, pop, nd, ng, nz, final):
    exp = FIX["population"][it]
    assert abs(pop - exp) <= TOL * abs(exp), "iteration %d population %.17g vs %.17g" % (it, pop, exp)
    for c in range(num_pr):
        got = canonical(get("phi", c), layout, M, ng, nz)
        assert rel_normwise(got, FIX["phi_it%d_pr%d" % (it, c)]) <= TOL, "phi it %d chunk %d" % (it, c)
    if final:
        for s in range(num_sdoms):
            got = canonical(get("psi", s), layout, nd, ng, nz)
            assert rel_normwise(got, FIX["psi_final_s%d" % s]) <= TOL, "psi chunk %d" % s


@pytest.mark.parametrize("kind", ["ref", "fma"])
@pytest.mark.parametrize("layout", list(LAYOUT_AXES))
def test_oracle_matches_golden(layout, kind):
    o = po.Oracle(po.make_input(**_cfg(layout)), kind=kind, omp=True)
    S0 = o.sdom(0)
    nd, ng, nz = S0["nd"], S0["ng"], S0["ni"] * S0["nj"] * S0["nk"]
    o.init_psi()
    for it in range(ITERS):
        _, comp = o.iterate()
        _check(o.get, o.num_pr, o.num_sdoms, layout, it, comp, nd, ng, nz, it == ITERS - 1)


def test_generator_matches_golden():
    o = po.Oracle(po.make_input(**_cfg("DGZ")), kind="ref", omp=False)
    assert np.array_equal(o.get("sigt_zonal", 0), FIX["sigt_zonal_pr0"])
    assert np.array_equal(o.get("volume", 0), FIX["volume_r0"])


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("layout", ["DGZ", "GZD", "ZDG"])
def test_product_matches_golden(ctx, layout, mode):
    import kripke_amd as ka
    hp = ka.HostProblem(ka.input_vars(**_cfg(layout)), ctx)
    o = po.Oracle(po.make_input(**_cfg(layout)), kind="ref", omp=False)
    S0 = o.sdom(0)
    nd, ng, nz = S0["nd"], S0["ng"], S0["ni"] * S0["nj"] * S0["nk"]
    if layout == "DGZ":
        assert np.array_equal(hp.get("sigt_zonal", 0), FIX["sigt_zonal_pr0"])
    hp.init_psi()
    for it in range(ITERS):
        pop = hp.iterate(mode)
        _check(hp.get, hp.num_pr, hp.num_sdoms, layout, it, pop, nd, ng, nz, it == ITERS - 1)
