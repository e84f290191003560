"""Shared helpers: build the same problem in the HIP product and in the CPU
oracle, and move field data between them."""
import numpy as np

import kripke_amd as ka
import pyoracle as po

# product field name -> oracle field name
ORACLE_NAME = {"quadrature/xcos": "xcos", "quadrature/ycos": "ycos", "quadrature/zcos": "zcos",
               "quadrature/w": "w", "quadrature/id": "id", "quadrature/jd": "jd", "quadrature/kd": "kd",
               "data/sigs": "sigs"}


def kwargs_for(layout="DGZ", **kw):
    base = dict(zones=(8, 8, 8), groups=8, quad=(4, 4), legendre=2, layout=layout, dset=8, gset=1,
                zset=(1, 1, 1), geometry=1, seed=20261202)
    base.update(kw)
    return base


def oracle_for(kind, **kw):
    return po.Oracle(po.make_input(**kw), kind=kind, omp=True)


def product_for(ctx, **kw):
    return ka.HostProblem(ka.input_vars(**kw), ctx)


def seeded_fill(hp, oracles, name, count, rng, scale=1.0, offset=0.0):
    """Write the same random values into `count` chunks of field `name`."""
    for c in range(count):
        n = hp.get(name, c).size
        v = offset + scale * rng.standard_normal(n)
        hp.set(name, c, v)
        for o in oracles:
            o.set(ORACLE_NAME.get(name, name), c, v)


def rel_normwise(a, b):
    d = np.max(np.abs(a - b)) if a.size else 0.0
    s = np.max(np.abs(b)) if b.size else 0.0
    return d / s if s > 0 else d


LAYOUT_AXES = {  # memory order of a (A, Group, Zone) field, slowest first
    "DGZ": "agz", "DZG": "azg", "GDZ": "gaz", "GZD": "gza", "ZDG": "zag", "ZGD": "zga"}


def canonical(arr, layout, na, ng, nz):
    """Field chunk in its nesting's memory order -> array indexed [a, g, z]."""
    order = LAYOUT_AXES[layout]
    ext = {"a": na, "g": ng, "z": nz}
    v = np.asarray(arr).reshape([ext[c] for c in order])
    return np.transpose(v, [order.index(c) for c in "agz"])


def phi_parity(hp, o, num_pr, layout, moments, ng, nz):
    """Worst, over PR chunks and moments, of the normwise relative error of
    each moment, and worst elementwise relative error with an absolute floor
    of 1e-12 times the moment's max-abs."""
    worst_norm, worst_elem = 0.0, 0.0
    for c in range(num_pr):
        a = canonical(hp.get("phi", c), layout, moments, ng, nz)
        b = canonical(o.get("phi", c), layout, moments, ng, nz)
        for nm in range(moments):
            x, y = a[nm], b[nm]
            scale = float(np.max(np.abs(y)))
            if scale == 0.0:
                worst_norm = max(worst_norm, float(np.max(np.abs(x))))
                continue
            worst_norm = max(worst_norm, float(np.max(np.abs(x - y))) / scale)
            floor = 1e-12 * scale
            worst_elem = max(worst_elem, float(np.max(np.abs(x - y) / np.maximum(np.abs(y), floor))))
    return worst_norm, worst_elem
