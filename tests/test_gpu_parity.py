"""Parity of the HIP path (through the C-ABI) against the CPU oracle.

Two oracle builds are used: 'fma' places fused multiply-adds exactly where
the HIP kernels do, so every kernel except the population reduction must
match it bit for bit; 'ref' has the reference's arithmetic (separate
multiply and add), which the HIP path must match to 1e-12 relative
(population, and phi normwise per moment), the bar of the task.
"""
import numpy as np
import pytest

from helpers import (canonical, kwargs_for, oracle_for, phi_parity, product_for, rel_normwise, seeded_fill)

LAYOUTS = ["DGZ", "DZG", "GDZ", "GZD", "ZDG", "ZGD"]
TOL_REF = 1e-12

pytestmark = pytest.mark.gpu


def _trio(ctx, **kw):
    return product_for(ctx, **kw), oracle_for("fma", **kw), oracle_for("ref", **kw)


# Zone-fastest nestings sweep a contiguous zone row with an affine scan of the
# x-face recurrence (re-associated arithmetic): their sweep outputs are held
# to a normwise bound against the fma-order oracle instead of bit equality.
ZFAST = ("DGZ", "GDZ")
TOL_SCAN = 1e-13


def _assert_chunks(hp, of, orf, name, count, tol=1e-13, exact=True):
    for c in range(count):
        a = hp.get(name, c)
        b = of.get(name, c)
        if exact:
            assert np.array_equal(a, b), "%s chunk %d differs from the fma oracle (max |d| %.3e)" % (
                name, c, float(np.max(np.abs(a - b))))
        else:
            assert rel_normwise(a, b) <= TOL_SCAN, "%s chunk %d: %.3e" % (name, c, rel_normwise(a, b))
        if orf is not None:
            assert rel_normwise(a, orf.get(name, c)) <= tol, name


@pytest.mark.parametrize("layout", LAYOUTS)
def test_ltimes(ctx, layout):
    kw = kwargs_for(layout, gset=2, zset=(2, 1, 1), groups=8)
    hp, of, orf = _trio(ctx, **kw)
    rng = np.random.default_rng(11)
    seeded_fill(hp, [of, orf], "psi", hp.num_sdoms, rng)
    seeded_fill(hp, [of, orf], "phi", hp.num_pr, rng)
    k = hp.kernels()
    k.ltimes(accumulate=True)
    of.ltimes()
    orf.ltimes()
    _assert_chunks(hp, of, orf, "phi", hp.num_pr)
    # overwrite form == kConst(phi, 0) followed by LTimes
    k.ltimes(accumulate=False)
    of.kconst("phi", 0.0)
    of.ltimes()
    _assert_chunks(hp, of, None, "phi", hp.num_pr)


@pytest.mark.parametrize("layout", ["DGZ", "GZD", "ZGD"])
def test_ltimes_lplustimes_many_directions(ctx, layout):
    """24 directions per direction set (GL 8 x 24): the direction-fastest
    kernels stage slabs wider than their 16-per-thread register prefetch."""
    kw = kwargs_for(layout, zset=(2, 1, 1), zones=(6, 4, 4), groups=4, quad=(8, 24))
    hp, of, orf = _trio(ctx, **kw)
    rng = np.random.default_rng(18)
    seeded_fill(hp, [of, orf], "psi", hp.num_sdoms, rng)
    seeded_fill(hp, [of, orf], "phi", hp.num_pr, rng)
    k = hp.kernels()
    k.ltimes(accumulate=True)
    of.ltimes()
    orf.ltimes()
    _assert_chunks(hp, of, orf, "phi", hp.num_pr)
    seeded_fill(hp, [of, orf], "phi_out", hp.num_pr, rng)
    k.lplustimes(accumulate=False)
    of.kconst("rhs", 0.0)
    of.lplustimes()
    _assert_chunks(hp, of, None, "rhs", hp.num_sdoms)


@pytest.mark.parametrize("layout", LAYOUTS)
def test_lplustimes(ctx, layout):
    kw = kwargs_for(layout, gset=2, zset=(1, 2, 1), groups=8)
    hp, of, orf = _trio(ctx, **kw)
    rng = np.random.default_rng(12)
    seeded_fill(hp, [of, orf], "phi_out", hp.num_pr, rng)
    seeded_fill(hp, [of, orf], "rhs", hp.num_sdoms, rng)
    k = hp.kernels()
    k.lplustimes(accumulate=True)
    of.lplustimes()
    orf.lplustimes()
    _assert_chunks(hp, of, orf, "rhs", hp.num_sdoms)
    k.lplustimes(accumulate=False)
    of.kconst("rhs", 0.0)
    of.lplustimes()
    _assert_chunks(hp, of, None, "rhs", hp.num_sdoms)


@pytest.mark.parametrize("layout", LAYOUTS)
def test_scattering_and_source(ctx, layout):
    kw = kwargs_for(layout, gset=2, zset=(1, 1, 2), groups=8)
    hp, of, orf = _trio(ctx, **kw)
    rng = np.random.default_rng(13)
    seeded_fill(hp, [of, orf], "phi", hp.num_pr, rng)
    seeded_fill(hp, [of, orf], "phi_out", hp.num_pr, rng)
    k = hp.kernels()
    k.scattering(accumulate=True)
    of.scattering()
    orf.scattering()
    _assert_chunks(hp, of, orf, "phi_out", hp.num_pr, tol=0.0)
    k.source(1.0)
    of.source()
    orf.source()
    _assert_chunks(hp, of, orf, "phi_out", hp.num_pr, tol=0.0)


@pytest.mark.parametrize("groups,legendre", [(4, 1), (40, 2)])
@pytest.mark.parametrize("layout", LAYOUTS)
def test_scattering_dense_sigs(ctx, layout, groups, legendre):
    """Group coupling and anisotropic moments: edits of data/sigs between
    calls switch the kernel to the dense path.  That path sums per material
    (sum_mat W(z,mat) sum_gp sigs phi) rather than per zone matrix, so it is
    held to 1e-13 normwise against the fma oracle and 1e-12 against the
    reference arithmetic.  40 groups in two sets: 20 per chunk, a partial
    output tile and a source-group remainder; the six nestings cover the
    zone-, group- and moment-fastest kernels."""
    kw = kwargs_for(layout, gset=2, groups=groups, legendre=legendre)
    hp, of, orf = _trio(ctx, **kw)
    rng = np.random.default_rng(14)
    sigs = rng.uniform(0.0, 0.1, hp.get("data/sigs").size)
    hp.set("data/sigs", 0, sigs)
    # the oracle holds sigs as [mat][n][g][gp]; convert from the nesting's order
    L1, G = legendre + 1, groups
    n_first = layout in ("DGZ", "DZG", "ZDG")
    s = sigs.reshape(3, L1, G, G) if n_first else np.transpose(sigs.reshape(3, G, G, L1), (0, 3, 1, 2))
    for o in (of, orf):
        o.set("sigs", 0, np.ascontiguousarray(s).ravel())
    seeded_fill(hp, [of, orf], "phi", hp.num_pr, rng)
    for o in (of, orf):
        o.kconst("phi_out", 0.0)
    hp.kernels().scattering(accumulate=False)
    of.scattering()
    orf.scattering()
    _assert_chunks(hp, of, orf, "phi_out", hp.num_pr, tol=TOL_REF, exact=False)
    # accumulate form adds onto what phi_out holds
    hp.kernels().scattering(accumulate=True)
    of.scattering()
    orf.scattering()
    _assert_chunks(hp, of, orf, "phi_out", hp.num_pr, tol=TOL_REF, exact=False)


@pytest.mark.parametrize("layout", LAYOUTS)
def test_sweep_subdomain(ctx, layout):
    """Kernel::sweepSubdomain on one subdomain: own planes in, own planes out."""
    kw = kwargs_for(layout, zset=(2, 2, 1), groups=4, zones=(8, 6, 10))
    hp, of, orf = _trio(ctx, **kw)
    rng = np.random.default_rng(15)
    seeded_fill(hp, [of, orf], "rhs", hp.num_sdoms, rng, scale=1.0, offset=2.0)
    for name in ("i_plane", "j_plane", "k_plane"):
        seeded_fill(hp, [of, orf], name, hp.num_sdoms, rng)
    k = hp.kernels()
    for s in (0, 5, hp.num_sdoms - 1):
        k.sweep_subdomain(s)
        of.sweep_subdomain(s)
        orf.sweep_subdomain(s)
    for name in ("psi", "i_plane", "j_plane", "k_plane"):
        _assert_chunks(hp, of, orf, name, hp.num_sdoms, tol=1e-13, exact=layout not in ZFAST)


@pytest.mark.parametrize("layout", LAYOUTS)
@pytest.mark.parametrize("zones,zset", [((8, 8, 8), (2, 2, 2)), ((12, 8, 6), (3, 2, 3))])
def test_sweep_solver_pipeline(ctx, layout, zones, zset):
    """On-device zone-set pipeline == SweepSolver's queue with plane hand-off."""
    kw = kwargs_for(layout, zset=zset, groups=8, zones=zones)
    hp, of, orf = _trio(ctx, **kw)
    rng = np.random.default_rng(16)
    seeded_fill(hp, [of, orf], "rhs", hp.num_sdoms, rng, offset=1.0)
    hp.kernels().sweep_solver(list(range(hp.num_sdoms)))
    of.sweep_solver()
    orf.sweep_solver()
    _assert_chunks(hp, of, orf, "psi", hp.num_sdoms, tol=1e-13, exact=layout not in ZFAST)
    # plane state after the pipeline (the reference leaves every subdomain's
    # outgoing faces in its own chunk, ParallelComm.cpp:156-162): subdomain t's
    # chunk along an axis holds t's outgoing faces where t is on the domain's
    # downwind boundary, its upwind neighbour's outgoing faces (t's incoming
    # faces) where it has neighbours on both sides, and is not written where
    # it has a downwind neighbour only
    checked = 0
    for t in range(hp.num_sdoms):
        info = of.sdom(t)
        for a, name in enumerate(("i_plane", "j_plane", "k_plane")):
            down, up = info["down%d" % a], info["up%d" % a]
            if down < 0:
                src = t
            elif up >= 0:
                src = up
            else:
                continue
            got, want = hp.get(name, t), of.get(name, src)
            checked += 1
            if layout in ZFAST:
                assert rel_normwise(got, want) <= TOL_SCAN, (t, name)
            else:
                assert np.array_equal(got, want), (t, name)
    assert checked >= hp.num_sdoms * 3 // 2


@pytest.mark.parametrize("layout", LAYOUTS)
def test_sweep_solver_block_jacobi(ctx, layout):
    """SweepSolver with block_jacobi (--pmethod bj): every subdomain swept
    with zero incoming faces.  psi and the outgoing faces left in each
    subdomain's planes against the oracle; the planes start from random
    values, which must not leak in."""
    kw = kwargs_for(layout, zset=(2, 2, 1), groups=4, zones=(8, 6, 10))
    hp, of, orf = _trio(ctx, **kw)
    rng = np.random.default_rng(24)
    seeded_fill(hp, [of, orf], "rhs", hp.num_sdoms, rng, offset=1.0)
    for name in ("i_plane", "j_plane", "k_plane"):
        seeded_fill(hp, [of, orf], name, hp.num_sdoms, rng)
    hp.kernels().sweep_solver(list(range(hp.num_sdoms)), block_jacobi=True)
    of.sweep_solver(block_jacobi=True)
    orf.sweep_solver(block_jacobi=True)
    for name in ("psi", "i_plane", "j_plane", "k_plane"):
        _assert_chunks(hp, of, orf, name, hp.num_sdoms, tol=1e-13, exact=layout not in ZFAST)


@pytest.mark.parametrize("layout", ["DGZ", "ZGD"])
def test_sweep_solver_block_jacobi_partial_list(ctx, layout):
    """A partial, unordered list: listed subdomains as in the full sweep,
    unlisted ones keep psi and end with zero planes (the old planes)."""
    kw = kwargs_for(layout, zset=(2, 2, 1), groups=4, zones=(8, 6, 10))
    hp, of, _ = _trio(ctx, **kw)
    rng = np.random.default_rng(25)
    seeded_fill(hp, [of], "rhs", hp.num_sdoms, rng, offset=1.0)
    seeded_fill(hp, [of], "psi", hp.num_sdoms, rng)
    for name in ("i_plane", "j_plane", "k_plane"):
        seeded_fill(hp, [of], name, hp.num_sdoms, rng)
    psi0 = [hp.get("psi", s) for s in range(hp.num_sdoms)]
    listed = [5, 0, 3]
    hp.kernels().sweep_solver(listed, block_jacobi=True)
    of.sweep_solver(block_jacobi=True)
    for s in range(hp.num_sdoms):
        if s in listed:
            a, b = hp.get("psi", s), of.get("psi", s)
            assert (np.array_equal(a, b) if layout not in ZFAST else rel_normwise(a, b) <= TOL_SCAN), s
        else:
            assert np.array_equal(hp.get("psi", s), psi0[s]), s
            for name in ("i_plane", "j_plane", "k_plane"):
                assert not np.any(hp.get(name, s)), (name, s)
    with pytest.raises(Exception):
        hp.kernels().sweep_solver([0, 0], block_jacobi=True)


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("layout", ["DGZ", "DZG", "GZD"])
def test_iterations_block_jacobi(ctx, layout, mode):
    """Source iterations with --pmethod bj (2x2x2 zone sets): population and
    phi per iteration against the oracle (1e-12), both modes."""
    kw = kwargs_for(layout, zones=(12, 8, 8), zset=(2, 2, 2), groups=8, quad=(4, 4), legendre=2, pmethod="bj")
    hp, of, orf = _trio(ctx, **kw)
    hp.init_psi()
    orf.init_psi()
    for it in range(4):
        pg = hp.iterate(mode)
        _, comp = orf.iterate()
        assert abs(pg - comp) <= TOL_REF * abs(comp), "iteration %d population %.17g vs %.17g" % (it, pg, comp)
        norm, _ = phi_parity(hp, orf, hp.num_pr, layout, 9, 8, 6 * 4 * 4)
        assert norm <= TOL_REF, "iteration %d phi normwise %.3e" % (it, norm)


@pytest.mark.parametrize("layout", ZFAST)
def test_sweep_kpatches(ctx, layout):
    """A 64 x 40 x 64 subdomain: 32 lanes per row (2 zones each) x 64 k-rows
    exceed one workgroup, so the row-scan sweep runs two k-patches per item
    and hands the k-face between them through the own plane."""
    kw = kwargs_for(layout, zones=(64, 40, 64), zset=(1, 1, 1), groups=2, quad=(4, 2), legendre=1)
    hp, of, orf = _trio(ctx, **kw)
    rng = np.random.default_rng(19)
    seeded_fill(hp, [of, orf], "rhs", hp.num_sdoms, rng, offset=1.0)
    hp.kernels().sweep_solver(list(range(hp.num_sdoms)))
    of.sweep_solver()
    orf.sweep_solver()
    _assert_chunks(hp, of, orf, "psi", hp.num_sdoms, tol=1e-13, exact=False)
    # Kernel::sweepSubdomain form (own planes in and out) on the same subdomains
    for s in range(hp.num_sdoms):
        hp.kernels().sweep_subdomain(s)
        of.sweep_subdomain(s)
    _assert_chunks(hp, of, None, "psi", hp.num_sdoms, exact=False)


# Row widths of the row-scan sweep: (zones, zone sets, GL quadrature); a row
# of ni zones takes W = ni/2 lanes (2 zones per lane).  W = 32 and 64 span
# several 16-lane DPP rows, where the scan crosses rows.
ROW_WIDTH_CASES = [
    ((32, 24, 16), (1, 1, 1), (4, 8)),   # W=16, one k-patch of 16 rows
    ((32, 12, 32), (1, 1, 1), (4, 8)),   # W=16, k-patches of 8 rows: k-face hand-off through the own plane
    ((16, 16, 16), (1, 2, 1), (8, 8)),   # W=8, pipeline with face hand-off
    ((64, 8, 8), (1, 1, 1), (4, 4)),     # W=32: two DPP rows per zone row
    ((128, 6, 8), (1, 1, 2), (4, 4)),    # W=64: four DPP rows per zone row, pipeline along k
]


@pytest.mark.parametrize("case", range(len(ROW_WIDTH_CASES)))
@pytest.mark.parametrize("layout", ZFAST)
def test_sweep_row_widths(ctx, layout, case):
    """Pipelined SweepSolver, the fused population and Kernel::sweepSubdomain
    for row widths of 8 to 64 lanes."""
    zones, zset, quad = ROW_WIDTH_CASES[case]
    kw = kwargs_for(layout, zones=zones, zset=zset, quad=quad, groups=2, legendre=1)
    hp, of, orf = _trio(ctx, **kw)
    rng = np.random.default_rng(26 + case)
    seeded_fill(hp, [of, orf], "rhs", hp.num_sdoms, rng, offset=1.0)
    k = hp.kernels()
    sdoms = list(range(hp.num_sdoms))
    k.sweep_solver(sdoms)
    of.sweep_solver()
    orf.sweep_solver()
    _assert_chunks(hp, of, orf, "psi", hp.num_sdoms, tol=1e-13, exact=False)
    pop = k.sweep_solver_population(sdoms)
    _assert_chunks(hp, of, None, "psi", hp.num_sdoms, exact=False)
    _, comp = orf.population()
    assert abs(pop - comp) <= TOL_REF * abs(comp)
    for name in ("i_plane", "j_plane", "k_plane"):
        seeded_fill(hp, [of], name, hp.num_sdoms, rng)
    for s in sdoms:
        k.sweep_subdomain(s)
        of.sweep_subdomain(s)
    for name in ("psi", "i_plane", "j_plane", "k_plane"):
        _assert_chunks(hp, of, None, name, hp.num_sdoms, exact=False)


@pytest.mark.parametrize("layout", ["DGZ", "DZG", "GZD"])
def test_sweep_many_directions(ctx, layout):
    """512 directions (64 per direction set) on a 32^3 subdomain: the sweep's
    coefficient table (64 x 97 doubles) plus the face buffers need more than
    64 KiB of LDS, which gfx950 allows per workgroup."""
    kw = kwargs_for(layout, zones=(32, 32, 32), zset=(1, 1, 1), dset=8, groups=1, quad=(16, 32), legendre=1)
    hp, of, orf = _trio(ctx, **kw)
    rng = np.random.default_rng(21)
    seeded_fill(hp, [of, orf], "rhs", hp.num_sdoms, rng, offset=1.0)
    hp.kernels().sweep_solver(list(range(hp.num_sdoms)))
    of.sweep_solver()
    orf.sweep_solver()
    _assert_chunks(hp, of, orf, "psi", hp.num_sdoms, tol=1e-13, exact=layout not in ZFAST)


@pytest.mark.parametrize("layout", ["DGZ", "ZGD"])
def test_population(ctx, layout):
    kw = kwargs_for(layout, zset=(1, 2, 1), gset=2)
    hp, of, orf = _trio(ctx, **kw)
    rng = np.random.default_rng(17)
    seeded_fill(hp, [of, orf], "psi", hp.num_sdoms, rng, offset=3.0)
    g = hp.kernels().population()
    lit, comp = orf.population()
    assert abs(g - comp) <= 1e-13 * abs(comp)
    assert g == hp.kernels().population(), "population must be deterministic"


@pytest.mark.parametrize("layout", ["DGZ", "GDZ", "DZG"])
@pytest.mark.parametrize("zones", [(8, 8, 8), (12, 6, 10)])
def test_sweep_solver_population(ctx, layout, zones):
    """SweepSolver + population in one call: psi bit-identical to
    kr_sweep_solver, population equal to kr_population to rounding and to
    the oracle's compensated sum; deterministic across calls."""
    kw = kwargs_for(layout, zset=(2, 2, 2), groups=8, zones=zones)
    hp, of, orf = _trio(ctx, **kw)
    rng = np.random.default_rng(23)
    seeded_fill(hp, [of, orf], "rhs", hp.num_sdoms, rng, offset=1.0)
    k = hp.kernels()
    sdoms = list(range(hp.num_sdoms))
    k.sweep_solver(sdoms)
    psi_a = [hp.get("psi", c) for c in range(hp.num_sdoms)]
    pop_a = k.population()
    pop_b = k.sweep_solver_population(sdoms[::-1])  # any complete list
    for c in range(hp.num_sdoms):
        assert np.array_equal(hp.get("psi", c), psi_a[c]), "psi chunk %d" % c
    assert abs(pop_b - pop_a) <= 1e-13 * abs(pop_a)
    orf.sweep_solver()
    _, comp = orf.population()
    assert abs(pop_b - comp) <= TOL_REF * abs(comp)
    assert pop_b == k.sweep_solver_population(sdoms), "fused population must be deterministic"
    with pytest.raises(Exception):
        k.sweep_solver_population(sdoms[:-1])


@pytest.mark.parametrize("layout", ["DGZ", "DZG"])
@pytest.mark.parametrize("uniform", [True, False])
def test_sweep_solver_population_volumes(ctx, layout, uniform):
    """The fused population with zone volumes that are all equal (the
    generator's, taken as one value per chunk) and that differ (read per
    zone): both against kr_population and the oracle on the same volumes,
    and again after the volumes change between calls."""
    kw = kwargs_for(layout, zset=(2, 1, 1), groups=8, zones=(8, 6, 10))
    hp, of, orf = _trio(ctx, **kw)
    rng = np.random.default_rng(29)
    seeded_fill(hp, [of, orf], "rhs", hp.num_sdoms, rng, offset=1.0)
    k = hp.kernels()
    sdoms = list(range(hp.num_sdoms))
    for trial in range(2):
        for c in range(hp.num_r):
            n = hp.get("volume", c).size
            v = np.full(n, 0.37 + trial) if uniform else rng.uniform(0.5, 2.0, n)
            hp.set("volume", c, v)
            orf.set("volume", c, v)
        pop = k.sweep_solver_population(sdoms)
        assert abs(pop - k.population()) <= 1e-13 * abs(pop)
        orf.sweep_solver()
        _, comp = orf.population()
        assert abs(pop - comp) <= TOL_REF * abs(comp), (trial, pop, comp)


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("layout", LAYOUTS)
def test_iterations_c0(ctx, layout, mode):
    """The CPU-runnable configuration (16^3 zones, 32 groups, 96 GL
    directions 8x12, P4, 8 direction sets, 10 iterations): per-iteration
    population within 1e-12 of the reference-arithmetic oracle's
    compensated sum, phi within 1e-12 normwise per moment, and phi/psi bit
    for bit equal to the fma-order oracle."""
    kw = kwargs_for(layout, zones=(16, 16, 16), groups=32, quad=(8, 12), legendre=4)
    hp, of, orf = _trio(ctx, **kw)
    hp.init_psi()
    of.init_psi()
    orf.init_psi()
    for it in range(10):
        pg = hp.iterate(mode)
        of.iterate()
        _, comp = orf.iterate()
        assert abs(pg - comp) <= TOL_REF * abs(comp), "iteration %d population %.17g vs %.17g" % (it, pg, comp)
        norm, elem = phi_parity(hp, orf, hp.num_pr, layout, 25, 32, 16 ** 3)
        assert norm <= TOL_REF, "iteration %d phi normwise %.3e" % (it, norm)
        # gross-error check per element (floor 1e-12 x the moment's max-abs):
        # FMA vs separate multiply-add rounding reaches ~1e-9 relative on
        # entries that cancel; the strict element check is the fma oracle below
        assert elem <= 1e-8, "iteration %d phi elementwise %.3e" % (it, elem)
        for c in range(hp.num_pr):
            a, b = hp.get("phi", c), of.get("phi", c)
            if layout in ZFAST:
                assert rel_normwise(a, b) <= TOL_SCAN * 10, "iteration %d phi chunk %d" % (it, c)
            else:
                assert np.array_equal(a, b), "iteration %d phi chunk %d" % (it, c)
    for s in range(hp.num_sdoms):
        a, b = hp.get("psi", s), of.get("psi", s)
        if layout in ZFAST:
            assert rel_normwise(a, b) <= TOL_SCAN * 10
        else:
            assert np.array_equal(a, b)


# Extents that the tuned shapes do not divide: (zones, zone sets, GL
# quadrature).  One-zone subdomains, odd rows (one zone per lane in the
# row-scan sweep), single-zone axes, one and three directions per direction
# set, and a 2x2x2 zone-set pipeline of 5x3x3 subdomains.
ODD_EXTENT_CASES = [
    ((1, 1, 1), (1, 1, 1), (4, 2)),
    ((3, 1, 2), (1, 1, 2), (4, 6)),
    ((10, 6, 6), (2, 2, 2), (4, 2)),
    ((7, 5, 3), (1, 1, 1), (4, 6)),
    ((1, 4, 3), (1, 2, 1), (4, 6)),
]


@pytest.mark.parametrize("case", range(len(ODD_EXTENT_CASES)))
@pytest.mark.parametrize("layout", LAYOUTS)
def test_iterations_odd_extents(ctx, layout, case):
    """Three source iterations from a seeded psi (so that scattering alone
    feeds the sweep where the source box misses the zones), alternating the
    reference call sequence and the on-device pipeline: population and phi
    against the reference-arithmetic oracle to 1e-12, phi and psi against
    the fma-order oracle.  With one zone along i the odd moments cancel by
    symmetry to rounding noise (~1e-17 against O(1) even moments), where a
    per-moment relative error divides noise by noise, so phi is held
    normwise per chunk there, and per moment wherever a subdomain has more
    than one zone along i."""
    zones, zset, quad = ODD_EXTENT_CASES[case]
    kw = kwargs_for(layout, zones=zones, zset=zset, groups=3, quad=quad, legendre=2)
    hp, of, orf = _trio(ctx, **kw)
    rng = np.random.default_rng(31 + case)
    seeded_fill(hp, [of, orf], "psi", hp.num_sdoms, rng, offset=1.0)
    for it in range(3):
        pg = hp.iterate(case % 2)
        of.iterate()
        _, comp = orf.iterate()
        assert comp != 0.0
        assert abs(pg - comp) <= TOL_REF * abs(comp), "iteration %d population %.17g vs %.17g" % (it, pg, comp)
        ni, nj, nk = (zones[a] // zset[a] for a in range(3))
        if ni > 1:
            norm, _ = phi_parity(hp, orf, hp.num_pr, layout, 9, 3, ni * nj * nk)
            assert norm <= TOL_REF, "iteration %d phi per-moment normwise %.3e" % (it, norm)
        for c in range(hp.num_pr):
            a, b = hp.get("phi", c), of.get("phi", c)
            norm = rel_normwise(a, orf.get("phi", c))
            assert norm <= TOL_REF, "iteration %d phi chunk %d normwise %.3e" % (it, c, norm)
            if layout in ZFAST:
                assert rel_normwise(a, b) <= TOL_SCAN * 10, "iteration %d phi chunk %d" % (it, c)
            else:
                assert np.array_equal(a, b), "iteration %d phi chunk %d" % (it, c)
    _assert_chunks(hp, of, orf, "psi", hp.num_sdoms, tol=1e-12, exact=layout not in ZFAST)


def test_modes_and_decomposition(ctx):
    """Reference call sequence (mode 0) and on-device pipeline (mode 1) give
    identical phi; zone-set / group-set decompositions agree to rounding."""
    base = kwargs_for("DGZ", zones=(8, 8, 8), groups=8)
    a = product_for(ctx, **base)
    b = product_for(ctx, **dict(base, zset=(2, 2, 2), gset=2))
    a.init_psi()
    b.init_psi()
    for _ in range(4):
        pa = a.iterate(1)
        pb = b.iterate(0)
        assert abs(pa - pb) <= 1e-13 * abs(pa)


def test_reference_default_problem(ctx):
    """The reference's own default run (Kobayashi-3i, dummy S2 with 96
    directions, 32 groups in 2 group sets, P4), 3 iterations."""
    kw = dict(zones=(16, 16, 16), groups=32, quad=None, directions=96, legendre=4, layout="DGZ", dset=8, gset=2,
              zset=(1, 1, 1), geometry=0)
    hp, of, orf = _trio(ctx, **kw)
    hp.init_psi()
    orf.init_psi()
    for it in range(3):
        pg = hp.iterate(1)
        _, comp = orf.iterate()
        assert abs(pg - comp) <= TOL_REF * abs(comp)


def test_nesting_invariance(ctx):
    """psi after two iterations is bit-identical across the nestings that share
    a sweep kernel (DZG/GZD/ZDG/ZGD: sequential recurrence; DGZ/GDZ: row scan),
    and within 1e-13 across the two families."""
    got = {}
    for layout in LAYOUTS:
        hp = product_for(ctx, **kwargs_for(layout, zset=(2, 1, 1), groups=8))
        hp.init_psi()
        for _ in range(2):
            hp.iterate(1)
        got[layout] = [canonical(hp.get("psi", s), layout, 2, 8, 4 * 8 * 8) for s in range(hp.num_sdoms)]
    for fam in (("DZG", "GZD", "ZDG", "ZGD"), ZFAST):
        for layout in fam[1:]:
            for x, y in zip(got[layout], got[fam[0]]):
                assert np.array_equal(x, y), layout
    for x, y in zip(got["DGZ"], got["ZDG"]):
        assert rel_normwise(x, y) <= TOL_SCAN


# --------------------------------------------------------------------------
# the autotuner's other sweep kernels for the nestings whose fastest index is
# a group or a direction, installed explicitly: the wave-uniform plane kernel
# and the systolic kernel with the faces between k-patches in LDS
# --------------------------------------------------------------------------
WAVE_LAYOUTS = ("DZG", "GZD", "ZDG", "ZGD")
# pairs per subdomain a multiple of 64, rows of 8/16/32 zones; (16, 8, 32)
# in 2 zone sets along i: 8 x 8 x 32 subdomains, two k-patches of 16 waves
WAVE_SHAPES = [dict(zones=(16, 8, 32), zset=(2, 1, 1), groups=32, quad=(4, 4)),
               dict(zones=(32, 16, 12), zset=(2, 2, 2), groups=64, quad=(4, 2))]


# the other sweep kernels the autotuner may install there: v_max 2 the wave
# kernel (kind 3), v_max 3 the systolic kernel with the LDS k-face ring (kind 4)
KERNELS = {"waves": (2, 3), "kring": (3, 4)}
# 32 x 16 x 32 subdomains with 4 directions per set: 4 j-rows x 16 k-layers
# x 16 pairs (group-fastest; four j-patches, two k-patches each), or every
# j-row x 16 k-layers x 4 pairs (direction-fastest; two k-patches), so the
# ring carries faces between k-patches in both patch orders; one and
# 2 x 1 x 2 zone sets
KRING_SHAPES = [dict(zones=(32, 16, 32), zset=(1, 1, 1), groups=64, quad=(4, 8)),
                dict(zones=(64, 16, 64), zset=(2, 1, 2), groups=32, quad=(4, 8))]
CASES = [("waves", WAVE_SHAPES[0]), ("waves", WAVE_SHAPES[1]), ("kring", KRING_SHAPES[0]),
         ("kring", KRING_SHAPES[1])]


def _install(hp, kernel):
    v, kind = KERNELS[kernel]
    got = hp.kernels().set_tuning(v_max=v)
    assert got == kind, "kernel %s does not cover this shape (installed kind %d)" % (kernel, got)


def _waves(hp):
    _install(hp, "waves")


@pytest.mark.parametrize("layout", WAVE_LAYOUTS)
@pytest.mark.parametrize("case", range(len(CASES)))
def test_wave_kernel_pipeline(ctx, layout, case):
    """Pipelined sweep with the fused population: psi bit for bit against the
    fma oracle, the population against the oracle's compensated sum, and the
    plane chunks on the domain's downwind boundary."""
    kernel, shape = CASES[case]
    kw = kwargs_for(layout, **shape)
    hp, of, orf = _trio(ctx, **kw)
    _install(hp, kernel)
    rng = np.random.default_rng(60 + case)
    seeded_fill(hp, [of, orf], "rhs", hp.num_sdoms, rng, offset=1.0)
    pop = hp.kernels().sweep_solver_population(list(range(hp.num_sdoms)))
    of.sweep_solver()
    orf.sweep_solver()
    _assert_chunks(hp, of, orf, "psi", hp.num_sdoms, tol=1e-13, exact=True)
    _, comp = orf.population()
    assert abs(pop - comp) <= TOL_REF * abs(comp)
    for t in range(hp.num_sdoms):
        info = of.sdom(t)
        for a, name in enumerate(("i_plane", "j_plane", "k_plane")):
            if info["down%d" % a] < 0:
                assert np.array_equal(hp.get(name, t), of.get(name, t)), (t, name)


@pytest.mark.parametrize("kernel", sorted(KERNELS))
@pytest.mark.parametrize("layout", WAVE_LAYOUTS)
def test_wave_kernel_block_jacobi_and_plain(ctx, layout, kernel):
    """Block Jacobi (every subdomain from zero incoming faces) and the plain
    pipelined sweep (no population) with the wave kernel, against the fma
    oracle bit for bit."""
    kw = kwargs_for(layout, **(WAVE_SHAPES[0] if kernel == "waves" else KRING_SHAPES[0]))
    hp, of, orf = _trio(ctx, **kw)
    _install(hp, kernel)
    rng = np.random.default_rng(64)
    seeded_fill(hp, [of, orf], "rhs", hp.num_sdoms, rng, offset=1.0)
    for name in ("i_plane", "j_plane", "k_plane"):
        seeded_fill(hp, [of, orf], name, hp.num_sdoms, rng)
    hp.kernels().sweep_solver(list(range(hp.num_sdoms)), block_jacobi=True)
    of.sweep_solver(block_jacobi=True)
    for name in ("psi", "i_plane", "j_plane", "k_plane"):
        _assert_chunks(hp, of, None, name, hp.num_sdoms, exact=True)
    seeded_fill(hp, [of], "rhs", hp.num_sdoms, rng, offset=1.0)
    hp.kernels().sweep_solver(list(range(hp.num_sdoms)))
    of.sweep_solver()
    _assert_chunks(hp, of, None, "psi", hp.num_sdoms, exact=True)


@pytest.mark.parametrize("kernel", sorted(KERNELS))
@pytest.mark.parametrize("layout", WAVE_LAYOUTS)
def test_wave_kernel_iterations(ctx, layout, kernel):
    """Three on-device source iterations with the wave kernel: phi and the
    populations equal those of the default (systolic) kernel bit for bit,
    and the populations agree with the reference-arithmetic oracle."""
    kw = kwargs_for(layout, **(WAVE_SHAPES[1] if kernel == "waves" else KRING_SHAPES[1]))
    a, b = product_for(ctx, **kw), product_for(ctx, **kw)
    _install(b, kernel)
    orf = oracle_for("ref", **kw)
    for x in (a, b, orf):
        x.init_psi()
    for it in range(3):
        pa, pb = a.iterate(1), b.iterate(1)
        _, comp = orf.iterate()
        assert abs(pb - comp) <= TOL_REF * abs(comp)
        assert abs(pa - pb) <= 1e-14 * abs(pa)
        for c in range(a.num_pr):
            assert np.array_equal(a.get("phi", c), b.get("phi", c)), (it, c)


def _vary_mesh(hp, oracles, zset, rng):
    """Zone widths that vary along every axis (the product holds one chunk
    per zone set along the axis, consecutive pieces of the oracle's global
    array) and zone volumes that vary per zone, written into the product and
    the oracles alike."""
    for ax, name in enumerate(("dx", "dy", "dz")):
        full = oracles[0].get(name)
        full = full * (0.75 + 0.5 * rng.random(full.size))
        for o in oracles:
            o.set(name, 0, full)
        n = full.size // zset[ax]
        for c in range(zset[ax]):
            hp.set(name, c, full[c * n:(c + 1) * n])
    for r in range(hp.num_r):
        v = hp.get("volume", r)
        v = v * (0.5 + rng.random(v.size))
        hp.set("volume", r, v)
        for o in oracles:
            o.set("volume", r, v)


@pytest.mark.parametrize("kernel", ["default", "kring"])
@pytest.mark.parametrize("layout", LAYOUTS)
def test_sweep_nonuniform_mesh(ctx, layout, kernel):
    """Pipelined sweep with the fused population on a mesh whose zone widths
    and volumes vary (the systolic kernel then reads its 2cos/delta tables
    per zone instead of one record per direction, and each zone's volume):
    psi bit for bit against the fma oracle (normwise for the row scan), the
    population against the oracle's compensated sum."""
    if kernel == "kring" and layout in ZFAST:
        pytest.skip("the k-ring variant serves the other four nestings")
    shape = KRING_SHAPES[0] if kernel == "kring" else dict(zones=(16, 8, 12), zset=(2, 2, 1), groups=32, quad=(4, 4))
    kw = kwargs_for(layout, **shape)
    hp, of, orf = _trio(ctx, **kw)
    if kernel == "kring":
        _install(hp, kernel)
    rng = np.random.default_rng(70)
    _vary_mesh(hp, [of, orf], shape["zset"], rng)
    seeded_fill(hp, [of, orf], "rhs", hp.num_sdoms, rng, offset=1.0)
    pop = hp.kernels().sweep_solver_population(list(range(hp.num_sdoms)))
    of.sweep_solver()
    orf.sweep_solver()
    _assert_chunks(hp, of, orf, "psi", hp.num_sdoms, tol=1e-13, exact=layout not in ZFAST)
    _, comp = orf.population()
    assert abs(pop - comp) <= TOL_REF * abs(comp)


# four zones per lane in the row-scan kernel (tuning point v_max = 4): rows of
# 32 zones (8 lanes), 28 (7 of 8 lanes live), 64 (16 lanes) and 8 (2 of 4
# lanes, eight zone sets)
ROWS4_SHAPES = [dict(zones=(32, 8, 12), zset=(1, 1, 1)), dict(zones=(28, 8, 8), zset=(1, 1, 1)),
                dict(zones=(64, 4, 8), zset=(1, 1, 1)), dict(zones=(16, 12, 20), zset=(2, 2, 2))]


@pytest.mark.parametrize("layout", ZFAST)
@pytest.mark.parametrize("shape", range(len(ROWS4_SHAPES)))
def test_row_scan_four_zones_per_lane(ctx, layout, shape):
    """Row-scan sweep with four zones per lane: the pipelined sweep with the
    fused population and the plain one against both oracles (normwise: the
    scan re-associates the row recurrence), psi different in its last bits
    from the two-zone kernel's (so the four-zone kernel ran), and three
    source iterations against the reference-arithmetic oracle."""
    kw = kwargs_for(layout, **ROWS4_SHAPES[shape])
    hp, of, orf = _trio(ctx, **kw)
    two = product_for(ctx, **kw)
    assert hp.kernels().set_tuning(v_max=4) == 1
    rng = np.random.default_rng(90 + shape)
    seeded_fill(hp, [of, orf], "rhs", hp.num_sdoms, rng, offset=1.0)
    for t in range(hp.num_sdoms):
        two.set("rhs", t, hp.get("rhs", t))
    pop = hp.kernels().sweep_solver_population(list(range(hp.num_sdoms)))
    two.kernels().sweep_solver_population(list(range(hp.num_sdoms)))
    of.sweep_solver()
    orf.sweep_solver()
    _assert_chunks(hp, of, orf, "psi", hp.num_sdoms, tol=1e-13, exact=False)
    _, comp = orf.population()
    assert abs(pop - comp) <= TOL_REF * abs(comp)
    assert any(not np.array_equal(hp.get("psi", t), two.get("psi", t)) for t in range(hp.num_sdoms))
    hp.kernels().sweep_solver(list(range(hp.num_sdoms)))
    _assert_chunks(hp, of, orf, "psi", hp.num_sdoms, tol=1e-13, exact=False)
    two.close()
    hp.close()
    b, ob = product_for(ctx, **kw), oracle_for("ref", **kw)
    assert b.kernels().set_tuning(v_max=4) == 1
    b.init_psi()
    ob.init_psi()
    for it in range(3):
        p = b.iterate(1)
        _, c = ob.iterate()
        assert abs(p - c) <= TOL_REF * abs(c), (it, p, c)
    b.close()


@pytest.mark.parametrize("layout", ZFAST)
def test_row_scan_four_zones_nonuniform_mesh(ctx, layout):
    """The row-scan sweep at four zones per lane on a mesh whose zone widths
    and volumes vary (each lane's four a values, the per-row volume loads):
    psi normwise against both oracles, the fused population against the
    oracle's compensated sum."""
    shape = dict(zones=(32, 8, 12), zset=(2, 2, 1), groups=32, quad=(4, 4))
    kw = kwargs_for(layout, **shape)
    hp, of, orf = _trio(ctx, **kw)
    assert hp.kernels().set_tuning(v_max=4) == 1
    rng = np.random.default_rng(71)
    _vary_mesh(hp, [of, orf], shape["zset"], rng)
    seeded_fill(hp, [of, orf], "rhs", hp.num_sdoms, rng, offset=1.0)
    pop = hp.kernels().sweep_solver_population(list(range(hp.num_sdoms)))
    of.sweep_solver()
    orf.sweep_solver()
    _assert_chunks(hp, of, orf, "psi", hp.num_sdoms, tol=1e-13, exact=False)
    _, comp = orf.population()
    assert abs(pop - comp) <= TOL_REF * abs(comp)
