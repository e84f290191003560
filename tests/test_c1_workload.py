"""Correctness at the production-share workload: 64^3 zones in 2x2x2 zone
sets, 128 groups, 128 Gauss-Legendre directions (8 polar x 16 azimuthal, 8
direction sets), P4 (25 moments), FP64 -- psi alone is 34.4 GB.

Three source iterations from psi = 0 (SteadyStateSolver.cpp:44-104) in each of
the six nestings, one problem alive on the device at a time:

- per-iteration populations agree across the nestings to 1e-12 relative;
- phi after the third iteration agrees bit for bit within each sweep-kernel
  family (zone-fastest row scan: DGZ, GDZ; systolic pencils: DZG, GZD, ZDG,
  ZGD) and normwise per moment to 1e-12 across the two families;
- DGZ in one zone set (a single 64^3 subdomain per direction set) agrees with
  the 2x2x2 pipeline to 1e-12;
- the fused population (2^32 terms summed by the sweep kernels as they write
  psi) matches a host sum of w(d)*psi*V(z) over the downloaded psi
  (Population.cpp:49-63) to 1e-12, and kr_population over the same psi;
- where the host has the memory (about 90 GB), the reference-arithmetic CPU
  oracle runs the same problem in DGZ for two iterations and the population
  and phi are held to it at 1e-12.

Each test prints its progress lines (flushed) so a long run shows where it is.
"""
import ctypes
import math
import os
import sys
import time

import numpy as np
import pytest

import kripke_amd as ka
from helpers import canonical, oracle_for

pytestmark = pytest.mark.gpu

C1 = dict(zones=(64, 64, 64), groups=128, quad=(8, 16), legendre=4, dset=8, gset=1, zset=(2, 2, 2), geometry=1,
          seed=20261202)
M, G, D = 25, 128, 128
ITERS = 3
TOL = 1e-12
ROW_FAMILY = ("DGZ", "GDZ")
PENCIL_FAMILY = ("DZG", "GZD", "ZDG", "ZGD")

_results = {}  # layout (or "DGZ-1zs") -> dict(pops=..., phi=[canonical per R chunk] or None)


def _say(msg):
    sys.stdout.write("[c1 %s] %s\n" % (time.strftime("%H:%M:%S"), msg))
    sys.stdout.flush()


def _download(ctx, ptr, n):
    out = np.empty(int(n), dtype=np.float64)
    ka.capi.check(ka.lib().kr_memcpy_dtoh(ctx.h, out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(ptr),
                                          int(n) * 8), "kr_memcpy_dtoh")
    return out


def _zone_sets(zset):
    nx, ny, nz = C1["zones"]
    return nx // zset[0], ny // zset[1], nz // zset[2]


def _global_phi(ctx, hp, layout, zset):
    """phi of every PR chunk (one group set) assembled into one [nm][g][k][j][i]
    array over the whole mesh.  PR chunk r = (zx*Sy + zy)*Sz + zz covers zones
    [zx*ni, ...) etc.; inside a chunk z = (k*nj + j)*ni + i."""
    ni, nj, nk = _zone_sets(zset)
    out = np.empty((M, G) + tuple(reversed(C1["zones"])), dtype=np.float64)
    for r in range(hp.num_pr):
        zz = r % zset[2]
        zy = (r // zset[2]) % zset[1]
        zx = r // (zset[2] * zset[1])
        ptr, n = hp.device_ptr("phi", r)
        c = canonical(_download(ctx, ptr, n), layout, M, G, ni * nj * nk).reshape(M, G, nk, nj, ni)
        out[:, :, zz * nk:(zz + 1) * nk, zy * nj:(zy + 1) * nj, zx * ni:(zx + 1) * ni] = c
    return out


def _run(ctx, layout, zset=(2, 2, 2), keep_phi=True, pop_pin=False):
    kw = dict(C1, zset=zset)
    t0 = time.time()
    hp = ka.HostProblem(ka.input_vars(layout=layout, **kw), ctx)
    try:
        hp.init_psi()
        pops = []
        for it in range(ITERS):
            pops.append(hp.iterate(1))
            _say("%s zset %s iteration %d population %.17g" % (layout, zset, it, pops[-1]))
        phi = _global_phi(ctx, hp, layout, zset) if keep_phi else None
        res = dict(pops=np.array(pops), phi=phi)
        if pop_pin:
            res["host_pop"], res["host_abs"] = _host_population(ctx, hp, layout, zset)
            res["kr_population"] = hp.kernels().population()
        _say("%s zset %s done in %.1f s" % (layout, zset, time.time() - t0))
        return res
    finally:
        hp.close()


def _host_population(ctx, hp, layout, zset):
    """sum over subdomains, d, g, z of w(d)*psi(d,g,z)*V(z) from the downloaded
    psi: each product rounded once in float64, summed pairwise per (subdomain,
    direction) slab of G*Z terms (numpy's pairwise summation), and the slab
    sums added exactly (math.fsum).  Also returns sum of |terms| (the scale of
    the rounding error)."""
    ni, nj, nk = _zone_sets(zset)
    nz = ni * nj * nk
    nd = D // C1["dset"]
    parts, abs_parts = [], []
    for s in range(hp.num_sdoms):
        ptr, n = hp.device_ptr("psi", s)
        psi = canonical(_download(ctx, ptr, n), layout, nd, G, nz)
        # direction set of subdomain s (SdomId = (gset*SQ + dset)*Sx*Sy*Sz + r)
        nr = zset[0] * zset[1] * zset[2]
        q, r = (s // nr) % C1["dset"], s % nr
        w = hp.get("quadrature/w", q)
        vol = hp.get("volume", r)
        for d in range(nd):
            t = (w[d] * psi[d]) * vol[None, :]
            parts.append(float(np.sum(t)))
            abs_parts.append(float(np.sum(np.abs(t))))
        del psi
    return math.fsum(parts), math.fsum(abs_parts)


def _get(ctx, key, layout, **kw):
    if key not in _results:
        _results[key] = _run(ctx, layout, **kw)
    return _results[key]


def _phi_rel_per_moment(a, b):
    worst = 0.0
    for nm in range(M):
        scale = float(np.max(np.abs(b[nm])))
        d = float(np.max(np.abs(a[nm] - b[nm])))
        worst = max(worst, d / scale if scale > 0 else d)
    return worst


@pytest.mark.timeout(900)
def test_c1_dgz_fused_population(ctx):
    """DGZ, the benchmarked nesting: per-iteration populations, and the fused
    population of the last iteration against the host sum over psi and
    against kr_population (a separate pass over psi)."""
    res = _get(ctx, "DGZ", "DGZ", pop_pin=True)
    pops = res["pops"]
    assert np.all(np.isfinite(pops)) and np.all(np.diff(pops) > 0), pops
    host, habs = res["host_pop"], res["host_abs"]
    rel = abs(pops[-1] - host) / abs(host)
    _say("fused population %.17g host %.17g rel %.3e (sum|terms|/|sum| %.3f); kr_population rel %.3e" % (
        pops[-1], host, rel, habs / abs(host), abs(res["kr_population"] - host) / abs(host)))
    assert rel <= TOL
    assert abs(res["kr_population"] - host) <= TOL * abs(host)


@pytest.mark.timeout(900)
@pytest.mark.parametrize("layout", ["GDZ", "DZG", "GZD", "ZDG", "ZGD"])
def test_c1_nesting(ctx, layout):
    """Every other nesting against DGZ (populations, phi per moment) and,
    inside its sweep-kernel family, against the family's first nesting bit
    for bit."""
    ref = _get(ctx, "DGZ", "DGZ", pop_pin=True)
    fam_first = ROW_FAMILY[0] if layout in ROW_FAMILY else PENCIL_FAMILY[0]
    res = _get(ctx, layout, layout)
    prel = np.max(np.abs(res["pops"] - ref["pops"]) / np.abs(ref["pops"]))
    phrel = _phi_rel_per_moment(res["phi"], ref["phi"])
    _say("%s vs DGZ: population rel %.3e, phi per-moment normwise %.3e" % (layout, prel, phrel))
    assert prel <= TOL
    assert phrel <= TOL
    if fam_first != layout:
        first = _get(ctx, fam_first, fam_first)
        # phi bit for bit; the populations are sums over psi in an order that
        # follows the nesting's own memory order, so they agree to rounding
        same = np.array_equal(res["phi"], first["phi"])
        _say("%s vs %s phi bit for bit: %s" % (layout, fam_first, same))
        assert same, "phi differs from %s in the same kernel family" % fam_first
    # the family's first nesting is needed by the later ones; drop the others' phi
    if layout not in (fam_first, "DGZ"):
        res["phi"] = None


@pytest.mark.timeout(900)
def test_c1_one_zone_set(ctx):
    """DGZ with one 64^3 subdomain per direction set (no zone-set pipeline)
    against the 2x2x2 pipeline: the same problem decomposed differently."""
    ref = _get(ctx, "DGZ", "DGZ", pop_pin=True)
    res = _run(ctx, "DGZ", zset=(1, 1, 1))
    prel = np.max(np.abs(res["pops"] - ref["pops"]) / np.abs(ref["pops"]))
    phrel = _phi_rel_per_moment(res["phi"], ref["phi"])
    _say("DGZ 1,1,1 vs 2,2,2: population rel %.3e, phi per-moment normwise %.3e" % (prel, phrel))
    assert prel <= TOL
    assert phrel <= TOL


def _host_memory_available():
    try:
        import psutil
        return psutil.virtual_memory().available
    except Exception:
        return 0


@pytest.mark.timeout(1100)
def test_c1_against_reference_arithmetic_oracle(ctx):
    """Two DGZ iterations of the same problem in the reference-arithmetic CPU
    oracle (OpenMP): population against its compensated sum, phi per moment
    normwise, both at 1e-12.  Needs about 90 GB of host memory for the
    oracle's fields; skipped with the reason where the host has less."""
    need = 100 * 2 ** 30
    avail = _host_memory_available()
    if avail < need:
        pytest.skip("host has %.0f GB available, the C1 oracle needs about 100 GB" % (avail / 2 ** 30))
    _results.pop("DZG", None)  # free the pencil family's phi before the oracle allocates
    t0 = time.time()
    orf = oracle_for("ref", layout="DGZ", **C1)
    orf.init_psi()
    _say("oracle created in %.1f s" % (time.time() - t0))
    hp = ka.HostProblem(ka.input_vars(layout="DGZ", **C1), ctx)
    try:
        hp.init_psi()
        ni, nj, nk = _zone_sets(C1["zset"])
        for it in range(2):
            pg = hp.iterate(1)
            t1 = time.time()
            _, comp = orf.iterate()
            rel = abs(pg - comp) / abs(comp)
            _say("oracle iteration %d in %.1f s: population %.17g vs %.17g rel %.3e" % (
                it, time.time() - t1, pg, comp, rel))
            assert rel <= TOL
            worst = 0.0
            for c in range(hp.num_pr):
                ptr, n = hp.device_ptr("phi", c)
                a = canonical(_download(ctx, ptr, n), "DGZ", M, G, ni * nj * nk)
                b = canonical(orf.get("phi", c), "DGZ", M, G, ni * nj * nk)
                worst = max(worst, _phi_rel_per_moment(a, b))
            _say("oracle iteration %d phi per-moment normwise %.3e" % (it, worst))
            assert worst <= TOL
    finally:
        hp.close()
        del orf
