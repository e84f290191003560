"""kr_sweep_solver_population_noalloc, the entry point that allocates nothing
(its partials live in a static device array), against
kr_sweep_solver_population, the entry point it is a variant of: psi and the
population bit for bit, on two calls in a row where the second problem is
smaller than the first, in both sweep-kernel families."""
import numpy as np
import pytest

import kripke_amd as ka
from helpers import kwargs_for

pytestmark = pytest.mark.gpu


def _pair(ctx, **kw):
    inp = ka.input_vars(**kw)
    return ka.HostProblem(inp, ctx), ka.HostProblem(inp, ctx)


def _fill_rhs(a, b, seed):
    rng = np.random.default_rng(seed)
    for s in range(a.num_sdoms):
        v = 1.0 + rng.random(a.get("rhs", s).size)
        a.set("rhs", s, v)
        b.set("rhs", s, v)


@pytest.mark.parametrize("layout", ["DGZ", "GZD"])
def test_noalloc_matches_sweep_solver_population(ctx, layout):
    shapes = [dict(zones=(16, 12, 16), zset=(2, 2, 2), groups=16), dict(zones=(8, 6, 8), zset=(2, 1, 2), groups=4)]
    first_units = None
    for call, shape in enumerate(shapes):
        a, b = _pair(ctx, **kwargs_for(layout, **shape))
        try:
            _fill_rhs(a, b, 70 + call)
            sdoms = list(range(a.num_sdoms))
            pa = a.kernels().sweep_solver_population(sdoms)
            pb = b.kernels().sweep_solver_population_noalloc(sdoms)
            assert pa == pb, (call, pa, pb)
            assert np.isfinite(pa) and pa > 0.0
            for s in sdoms:
                assert np.array_equal(a.get("psi", s), b.get("psi", s)), (call, s)
            # the same call again on the same problem (the scratch keeps the last call's contents)
            assert b.kernels().sweep_solver_population_noalloc(sdoms) == pa
            n = a.num_sdoms * a.get("psi", 0).size
            if first_units is None:
                first_units = n
            else:
                assert n < first_units  # the second call's problem is the smaller one
        finally:
            a.close()
            b.close()


def test_noalloc_rejects_partial_list(ctx):
    a = ka.HostProblem(ka.input_vars(**kwargs_for("DGZ", zones=(8, 6, 8), zset=(2, 1, 2), groups=4)), ctx)
    try:
        with pytest.raises(ka.KripkeError):
            a.kernels().sweep_solver_population_noalloc([0])
        with pytest.raises(ka.KripkeError):
            a.kernels().sweep_solver_population_noalloc([0] * a.num_sdoms)
    finally:
        a.close()
