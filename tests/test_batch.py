"""Batched persistent launch (kr_batch): every problem of a batch gives the
same populations and phi as its own iterations run one call at a time, bit
for bit, across several graph replays with eager work in between and a
re-capture when the iteration count changes."""
import numpy as np
import pytest

import kripke_amd as ka

pytestmark = pytest.mark.gpu


def mixed_problems(n, seed0=100):
    """Small problems of varied shape, nesting, groups, directions and
    scattering order (the batched configuration, scaled down)."""
    rng = np.random.default_rng(seed0)
    layouts = ["DGZ", "DZG", "GDZ", "GZD", "ZDG", "ZGD"]
    out = []
    for i in range(n):
        zones = tuple(int(v) for v in np.clip(8 + rng.integers(-2, 5, size=3), 6, 12))
        out.append(dict(zones=zones, groups=int(rng.choice([4, 8])), quad=(4, int(rng.choice([2, 4]))),
                        legendre=int(rng.integers(1, 4)), layout=layouts[i % 6], dset=8, gset=1,
                        zset=(2, 1, 1) if (i % 2 and zones[0] % 2 == 0) else (1, 1, 1), geometry=1,
                        seed=seed0 + i))
    return out


def test_batch_matches_single_problem_runs(ctx):
    cfgs = mixed_problems(6)
    ref_pops, ref_phi = [], []
    for c in cfgs:
        hp = ka.HostProblem(ka.input_vars(**c), ctx)
        hp.init_psi()
        ref_pops.append([hp.iterate(1) for _ in range(5)])
        ref_phi.append([hp.get("phi", k) for k in range(hp.num_pr)])
        hp.close()
    hps = [ka.HostProblem(ka.input_vars(**c), ctx) for c in cfgs]
    for hp in hps:
        hp.init_psi()
    b = ka.Batch(ctx, hps)
    p1 = b.iterate(2)
    hps[0].kernels().population()          # eager work between replays
    p2 = b.iterate(2)
    ctx.synchronize()
    p3 = b.iterate(1)                       # re-capture with another iteration count
    got = np.concatenate([p1, p2, p3], axis=1)
    for i in range(len(cfgs)):
        assert np.array_equal(got[i], np.array(ref_pops[i])), "problem %d: %s vs %s" % (i, got[i], ref_pops[i])
        for k in range(hps[i].num_pr):
            assert np.array_equal(hps[i].get("phi", k), ref_phi[i][k])
    b.close()
    for hp in hps:
        hp.close()


def zone_fastest_problems(n, seed0=300):
    """DGZ and GDZ problems whose extents differ inside one row-scan launch:
    odd and even ni (zones per lane 1 or 2, row groups of 8 to 32 lanes) and
    k extents 5 to 13 (k-rows per block), one or two zone sets along x."""
    rng = np.random.default_rng(seed0)
    out = []
    for i in range(n):
        zones = (int(rng.integers(5, 14)) * (2 if i % 3 == 0 else 1), int(rng.integers(6, 12)),
                 int(rng.integers(5, 14)))
        out.append(dict(zones=zones, groups=int(rng.choice([4, 8])), quad=(4, int(rng.choice([2, 4]))),
                        legendre=int(rng.integers(1, 4)), layout="DGZ" if i % 2 == 0 else "GDZ", dset=8, gset=1,
                        zset=(2, 1, 1) if i % 3 == 0 else (1, 1, 1), geometry=1, seed=seed0 + i))
    return out


def test_batch_mixed_extents_and_coupled_sigs(ctx):
    """Multi-problem launches over subdomains of different extents and the
    dense scattering of problems whose sigs couples groups: populations and
    phi equal each problem's own runs bit for bit, also after a problem's sigs
    changes between calls (a re-capture)."""
    cfgs = zone_fastest_problems(9)
    rng = np.random.default_rng(5)
    coupled = {}
    for i in (0, 4):
        hp = ka.HostProblem(ka.input_vars(**cfgs[i]), ctx)
        coupled[i] = rng.uniform(0.0, 0.05, hp.get("data/sigs").size)
        hp.close()

    def run(hps, step):
        hps[0].set("data/sigs", 0, coupled[0])
        a = step(3)
        hps[4].set("data/sigs", 0, coupled[4])
        return np.concatenate([a, step(2)], axis=1)

    ref = [ka.HostProblem(ka.input_vars(**c), ctx) for c in cfgs]
    for hp in ref:
        hp.init_psi()
    want = run(ref, lambda k: np.array([[hp.iterate(1) for _ in range(k)] for hp in ref]))
    hps = [ka.HostProblem(ka.input_vars(**c), ctx) for c in cfgs]
    for hp in hps:
        hp.init_psi()
    b = ka.Batch(ctx, hps)
    got = run(hps, b.iterate)
    for i in range(len(cfgs)):
        assert np.array_equal(got[i], want[i]), "problem %d (%s): %s vs %s" % (i, cfgs[i], got[i], want[i])
        for k in range(hps[i].num_pr):
            assert np.array_equal(hps[i].get("phi", k), ref[i].get("phi", k)), "problem %d chunk %d" % (i, k)
        for k in range(hps[i].num_sdoms):
            assert np.array_equal(hps[i].get("psi", k), ref[i].get("psi", k)), "problem %d sdom %d" % (i, k)
    b.close()
    for hp in hps + ref:
        hp.close()


# (zones, zone sets, GL quadrature) that the tuned shapes do not divide:
# one-zone subdomains, odd rows, single-zone axes, one and three directions
# per direction set, and a 2x2x2 pipeline of 5x3x3 subdomains.
ODD_EXTENTS = [((1, 1, 1), (1, 1, 1), (4, 2)), ((3, 1, 2), (1, 1, 2), (4, 6)), ((10, 6, 6), (2, 2, 2), (4, 2)),
               ((7, 5, 3), (1, 1, 1), (4, 6)), ((1, 4, 3), (1, 2, 1), (4, 6))]


def test_batch_odd_extents(ctx):
    """Problems of those extents in all six nestings in one batch, from a
    seeded psi (so that scattering feeds the sweep where the source box
    misses the zones): populations, phi and psi equal each problem's own
    runs bit for bit."""
    layouts = ["DGZ", "DZG", "GDZ", "GZD", "ZDG", "ZGD"]
    cfgs = [dict(zones=z, zset=zs, quad=q, groups=3, legendre=2, layout=layouts[i % 6], dset=8, gset=1, geometry=1,
                 seed=500 + i) for i, (z, zs, q) in enumerate(ODD_EXTENTS * 2)]

    def seeded(hps):
        rng = np.random.default_rng(9)
        for hp in hps:
            for s in range(hp.num_sdoms):
                hp.set("psi", s, 1.0 + rng.standard_normal(hp.get("psi", s).size))
        return hps

    ref = seeded([ka.HostProblem(ka.input_vars(**c), ctx) for c in cfgs])
    want = np.array([[hp.iterate(1) for _ in range(3)] for hp in ref])
    hps = seeded([ka.HostProblem(ka.input_vars(**c), ctx) for c in cfgs])
    b = ka.Batch(ctx, hps)
    got = b.iterate(3)
    for i in range(len(cfgs)):
        assert np.all(want[i] != 0.0)
        assert np.array_equal(got[i], want[i]), "problem %d (%s): %s vs %s" % (i, cfgs[i], got[i], want[i])
        for k in range(hps[i].num_pr):
            assert np.array_equal(hps[i].get("phi", k), ref[i].get("phi", k)), "problem %d chunk %d" % (i, k)
        for k in range(hps[i].num_sdoms):
            assert np.array_equal(hps[i].get("psi", k), ref[i].get("psi", k)), "problem %d sdom %d" % (i, k)
    b.close()
    for hp in hps + ref:
        hp.close()


def test_batch_rejects_bad_input(ctx):
    hp = ka.HostProblem(ka.input_vars(**mixed_problems(1)[0]), ctx)
    with pytest.raises(ka.KripkeError):
        ka.Batch(ctx, [hp, hp])
    hp.close()


@pytest.mark.parametrize("layout", ["DGZ", "DZG"])
def test_autotune_keeps_results(ctx, layout):
    """The tuned sweep schedule changes blocking only: populations and phi
    agree with the untuned problem (scan re-association at most), and a
    second problem of the same shape reuses the cached choice."""
    cfg = dict(zones=(16, 16, 16), groups=8, quad=(4, 4), legendre=2, layout=layout, dset=8, gset=1,
               zset=(2, 2, 2), geometry=1, seed=7)
    a = ka.HostProblem(ka.input_vars(**cfg), ctx)
    b = ka.HostProblem(ka.input_vars(**cfg), ctx)
    choice = b.kernels().autotune()
    assert choice[2] in (50, 100, 200)
    assert b.kernels().autotune() == choice  # cached
    c = ka.HostProblem(ka.input_vars(**cfg), ctx)
    assert c.kernels().autotune() == choice
    for hp in (a, b):
        hp.init_psi()
    for _ in range(3):
        pa, pb = a.iterate(1), b.iterate(1)
        assert abs(pa - pb) <= 1e-13 * abs(pa)
    for k in range(a.num_pr):
        x, y = a.get("phi", k), b.get("phi", k)
        assert np.max(np.abs(x - y)) <= 1e-12 * np.max(np.abs(x))
    # a batch built after tuning, and re-captured after re-tuning, still matches
    bt = ka.Batch(ctx, [c])
    c.init_psi()
    pops = bt.iterate(3)[0]
    a.init_psi()
    ref = [a.iterate(1) for _ in range(3)]
    assert np.allclose(pops, ref, rtol=1e-13, atol=0)
    bt.close()
    for hp in (a, b, c):
        hp.close()


def test_autotune_nesting(ctx):
    """The nesting autotuner returns one of the six nestings, caches it per
    shape, and a problem built in that nesting iterates like the DGZ one
    (populations within the row-scan re-association)."""
    cfg = dict(zones=(10, 9, 8), groups=8, quad=(4, 2), legendre=2, layout="DGZ", dset=8, gset=1, zset=(1, 1, 1),
               geometry=1, seed=11)
    lay, ms = ka.autotune_nesting(ctx, ka.input_vars(**cfg))
    assert lay in ka.LAYOUTS and ms > 0
    assert ka.autotune_nesting(ctx, ka.input_vars(**dict(cfg, seed=12))) == (lay, ms)  # cached per shape
    a = ka.HostProblem(ka.input_vars(**cfg), ctx)
    b = ka.HostProblem(ka.input_vars(**dict(cfg, layout=lay)), ctx)
    for hp in (a, b):
        hp.init_psi()
    for _ in range(3):
        pa, pb = a.iterate(1), b.iterate(1)
        assert abs(pa - pb) <= 1e-12 * abs(pa)
    a.close()
    b.close()


def test_batch_refuses_after_member_destroyed(ctx):
    """A member problem destroyed while its batch lives: kr_batch_iterate
    fails with KR_ERR_INVALID (no launch over freed tables), and the batch
    is still released cleanly."""
    cfgs = mixed_problems(2, seed0=700)
    hps = [ka.HostProblem(ka.input_vars(**c), ctx) for c in cfgs]
    for hp in hps:
        hp.init_psi()
    b = ka.Batch(ctx, hps)
    b.iterate(1)
    hps[1].close()
    with pytest.raises(ka.KripkeError) as e:
        b.iterate(1)
    assert e.value.code == 1 and "destroyed" in str(e.value)
    b.close()
    # the surviving problem still runs on its own
    assert np.isfinite(hps[0].iterate(1))
    hps[0].close()


def test_batch_systolic_tuning_points(ctx):
    """Problems of the group- and direction-fastest nestings with the k-ring
    variant (patches in k-fastest order, faces between k-patches in LDS) and
    the wave-uniform kernel installed, two of them with varying zone widths
    (the systolic kernel's instance for whole table rows) in one batch:
    populations and phi equal each problem's own runs bit for bit."""
    base = dict(zones=(32, 16, 32), zset=(1, 1, 1), groups=64, quad=(4, 8), legendre=2, dset=8, gset=1, geometry=1)
    cfgs = [dict(base, layout=lay, seed=700 + i) for i, lay in enumerate(["DZG", "GZD", "ZDG", "ZGD", "DZG", "GZD"])]
    tuning = [3, 3, 3, 3, 2, 0]  # k-ring x4, wave kernel, plain systolic
    widths = {1: np.linspace(0.8, 1.2, 32), 4: np.linspace(1.3, 0.7, 32)}

    kinds = [None] * len(cfgs)

    def make(i):
        hp = ka.HostProblem(ka.input_vars(**cfgs[i]), ctx)
        if i in widths:  # zone widths along x vary: the tables' rows are not uniform
            dx = hp.get("dx", 0)
            hp.set("dx", 0, dx * widths[i])
        kinds[i] = hp.kernels().set_tuning(v_max=tuning[i])
        hp.init_psi()
        return hp

    ref_pops, ref_phi = [], []
    for i in range(len(cfgs)):
        hp = make(i)
        ref_pops.append([hp.iterate(1) for _ in range(3)])
        ref_phi.append([hp.get("phi", k) for k in range(hp.num_pr)])
        hp.close()
    hps = [make(i) for i in range(len(cfgs))]
    assert kinds == [4, 4, 4, 4, 3, 0], kinds
    b = ka.Batch(ctx, hps)
    p1 = b.iterate(2)
    p2 = b.iterate(1)
    got = np.concatenate([p1, p2], axis=1)
    for i in range(len(cfgs)):
        assert np.array_equal(got[i], np.array(ref_pops[i])), "problem %d: %s vs %s" % (i, got[i], ref_pops[i])
        for k in range(hps[i].num_pr):
            assert np.array_equal(hps[i].get("phi", k), ref_phi[i][k])
    b.close()
    for hp in hps:
        hp.close()



def test_batch_row_scan_four_zones_per_lane(ctx):
    """Zone-fastest problems with the row-scan kernel at four zones per lane
    installed (rows of 32 and 16 zones, one mesh with varying zone widths)
    beside one at the default two, in one batch: populations and phi equal
    each problem's own runs bit for bit."""
    base = dict(zset=(1, 1, 1), groups=32, quad=(4, 8), legendre=2, dset=8, gset=1, geometry=1)
    cfgs = [dict(base, zones=(32, 16, 12), layout="DGZ", seed=800), dict(base, zones=(16, 8, 8), layout="GDZ", seed=801),
            dict(base, zones=(32, 8, 16), layout="GDZ", seed=802), dict(base, zones=(32, 16, 12), layout="DGZ", seed=803)]
    tuning = [4, 4, 4, 0]
    widths = {2: np.linspace(0.8, 1.2, 32)}

    def make(i):
        hp = ka.HostProblem(ka.input_vars(**cfgs[i]), ctx)
        if i in widths:
            dx = hp.get("dx", 0)
            hp.set("dx", 0, dx * widths[i])
        assert hp.kernels().set_tuning(v_max=tuning[i]) == 1
        hp.init_psi()
        return hp

    ref_pops, ref_phi = [], []
    for i in range(len(cfgs)):
        hp = make(i)
        ref_pops.append([hp.iterate(1) for _ in range(3)])
        ref_phi.append([hp.get("phi", k) for k in range(hp.num_pr)])
        hp.close()
    hps = [make(i) for i in range(len(cfgs))]
    b = ka.Batch(ctx, hps)
    got = np.concatenate([b.iterate(2), b.iterate(1)], axis=1)
    for i in range(len(cfgs)):
        assert np.array_equal(got[i], np.array(ref_pops[i])), "problem %d: %s vs %s" % (i, got[i], ref_pops[i])
        for k in range(hps[i].num_pr):
            assert np.array_equal(hps[i].get("phi", k), ref_phi[i][k])
    b.close()
    for hp in hps:
        hp.close()
