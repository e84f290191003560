"""Independent pins of the oracle's input assembly (no GPU).

Covers the three places where the generated inputs are assembled from the
geometry, which the product's generator shares with the oracle only by being
bitwise equal to it:

- the mixed-element fractions of every zone from 4^3 sub-samples at
  (s+1)/(n_s+1) of the zone (src/Kripke/Generate/Space.cpp:189-236) and the
  mixed-element tables built from them (Space.cpp:255-310);
- sigt_zonal(g,z) = sum over the zone's mixed elements of
  fraction * sigma_t[material] (Space.cpp:334-363);
- data/sigs = sigma_s[material] on the group diagonal of Legendre order 0,
  zero elsewhere (src/Kripke/Generate/Data.cpp:87-101).

The fractions are recomputed here a different way: the boxes are
axis-aligned, so the number of sub-samples of zone (i,j,k) that fall inside a
box is the product of three one-dimensional counts, and material 1 (its box
minus the material-0 box, which is painted last) is the count in its box less
the count in the intersection of the two boxes.  The box corners and cross
sections come from an independent Python SplitMix64.  The Kobayashi-3i
geometry is checked by evaluating its region table on the whole sub-sample
grid at once and by the exact region volumes on a mesh whose zone faces lie
on the region boundaries.
"""
import numpy as np
import pytest

import kripke_amd as ka
import pyoracle as po
from helpers import canonical

LAYOUTS = ["DGZ", "DZG", "GDZ", "GZD", "ZDG", "ZGD"]


def _splitmix64(seed, n):
    mask = (1 << 64) - 1
    s, out = seed, []
    for _ in range(n):
        s = (s + 0x9E3779B97F4A7C15) & mask
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & mask
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & mask
        out.append(z ^ (z >> 31))
    return out


def seeded_inputs(seed, edge=100.0):
    """Boxes (material 1's, then material 0's) and per-material sigma_t,
    sigma_s from the stated generator: side = edge*(0.1 + 0.4u),
    lo = (edge - side)u per axis; sigma_t = 10^(-4 + 4u); sigma_s =
    (0.3 + 0.6u) sigma_t."""
    u = [(x >> 11) * 2.0 ** -53 for x in _splitmix64(seed, 18)]
    boxes = []
    for b in range(2):
        lo, hi = [], []
        for ax in range(3):
            side = edge * (0.1 + 0.4 * u[b * 6 + 2 * ax])
            l0 = (edge - side) * u[b * 6 + 2 * ax + 1]
            lo.append(l0)
            hi.append(l0 + side)
        boxes.append((np.array(lo), np.array(hi)))
    sigt = np.array([10.0 ** (-4.0 + 4.0 * u[12 + m]) for m in range(3)])
    sigs = np.array([(0.3 + 0.6 * u[15 + m]) * sigt[m] for m in range(3)])
    return boxes, sigt, sigs


def axis_samples(x0, dx, n, ns):
    """Sub-sample coordinates along one axis of a chunk starting at x0 with n
    zones: [zone, sample], the zone origin x0 + dx*i plus dx*(s+1)/(ns+1)."""
    i = np.arange(n, dtype=np.float64)[:, None]
    s = np.arange(ns, dtype=np.float64)[None, :]
    return (x0 + dx * i) + dx * (s + 1) / (ns + 1)


def box_counts(samples, lo, hi):
    """[k, j, i] number of sub-samples inside the closed box, as the product
    of the per-axis counts (samples = (x, y, z) arrays [zone, sample])."""
    c = [np.sum((samples[a] >= lo[a]) & (samples[a] <= hi[a]), axis=1) for a in range(3)]
    return c[2][:, None, None] * c[1][None, :, None] * c[0][None, None, :]


def expected_mix_seeded(kw, r_chunk, boxes):
    """Fractions [zone, material] of R chunk r_chunk (zone = (k*nj + j)*ni + i)."""
    nx, ny, nz = kw["zones"]
    sx, sy, sz = kw["zset"]
    ns = kw.get("subsamples", 4)
    ni, nj, nk = nx // sx, ny // sy, nz // sz
    zz, zy, zx = r_chunk % sz, (r_chunk // sz) % sy, r_chunk // (sz * sy)
    dx, dy, dz = 100.0 / nx, 100.0 / ny, 100.0 / nz
    samples = (axis_samples(0.0 + dx * (zx * ni), dx, ni, ns), axis_samples(0.0 + dy * (zy * nj), dy, nj, ns),
               axis_samples(0.0 + dz * (zz * nk), dz, nk, ns))
    (lo1, hi1), (lo0, hi0) = boxes
    n0 = box_counts(samples, lo0, hi0)
    lo01, hi01 = np.maximum(lo0, lo1), np.minimum(hi0, hi1)
    n01 = box_counts(samples, lo01, hi01) if np.all(lo01 <= hi01) else 0
    n1 = box_counts(samples, lo1, hi1) - n01
    n2 = ns ** 3 - n0 - n1
    cnt = np.stack([n0.ravel(), n1.ravel(), n2.ravel()], axis=1)
    assert np.all(cnt >= 0)
    return cnt / float(ns ** 3)


def mix_tables(frac):
    """zone_to_mixelem, zone_to_num_mixelem, mixelem_to_zone/material/fraction
    of Space.cpp:282-310: per zone, the materials with a non-zero fraction in
    ascending order."""
    nzn = frac.shape[0]
    present = frac > 0.0
    num = present.sum(axis=1)
    first = np.concatenate([[0], np.cumsum(num)[:-1]])
    zone, mat = np.nonzero(present)  # row-major: zones ascending, materials ascending
    return first, num, zone, mat, frac[zone, mat], nzn


def expected_sigt(frac, sigt_mat):
    """sigt(z) accumulated over the zone's mixed elements in material order
    (absent materials add +0.0, which changes nothing)."""
    acc = np.zeros(frac.shape[0])
    for m in range(3):
        acc = acc + frac[:, m] * sigt_mat[m]
    return acc


SEEDED_CASES = [
    dict(zones=(12, 10, 9), zset=(2, 1, 3), groups=3, seed=20261202),
    dict(zones=(16, 16, 16), zset=(2, 2, 2), groups=2, seed=1),
    dict(zones=(7, 11, 5), zset=(1, 1, 1), groups=1, seed=12345),
    dict(zones=(20, 8, 12), zset=(4, 2, 3), groups=2, seed=987654321),
]


def _oracle(kw, layout="DGZ"):
    return po.Oracle(po.make_input(quad=(4, 4), legendre=1, dset=8, gset=1, geometry=1, layout=layout,
                                   **{k: v for k, v in kw.items()}))


@pytest.mark.parametrize("case", range(len(SEEDED_CASES)))
def test_mixed_elements_seeded(case):
    kw = SEEDED_CASES[case]
    boxes, _, _ = seeded_inputs(kw["seed"])
    o = _oracle(kw)
    nr = kw["zset"][0] * kw["zset"][1] * kw["zset"][2]
    vols = np.zeros(3)
    for r in range(nr):
        frac = expected_mix_seeded(kw, r, boxes)
        first, num, zone, mat, fr, nzn = mix_tables(frac)
        assert np.array_equal(o.get("zone_to_mixelem", r), first)
        assert np.array_equal(o.get("zone_to_num_mixelem", r), num)
        assert np.array_equal(o.get("mixelem_to_zone", r), zone)
        assert np.array_equal(o.get("mixelem_to_material", r), mat)
        assert np.array_equal(o.get("mixelem_to_fraction", r), fr)
        vol = o.get("volume", r)
        assert np.all(vol == vol[0])
        vols += frac.sum(axis=0) * vol[0]
    # the sampled material volumes approximate the boxes' exact volumes: each
    # sample stands for dx/4 along its axis, so each box extent is off by at
    # most one zone width per axis
    (lo1, hi1), (lo0, hi0) = boxes
    d = 100.0 / np.array(kw["zones"], dtype=float)
    v0 = np.prod(hi0 - lo0)
    bound0 = np.prod(hi0 - lo0 + 2 * d) - v0
    assert abs(vols[0] - v0) <= bound0
    assert abs(vols.sum() - 1e6) <= 1e-9 * 1e6


@pytest.mark.parametrize("layout", LAYOUTS)
@pytest.mark.parametrize("case", [0, 3])
def test_sigt_zonal_and_sigs_seeded(case, layout):
    """sigt_zonal from the recomputed fractions and the independently drawn
    sigma_t (not the oracle's), in every nesting's layout; sigs from the
    independently drawn sigma_s."""
    kw = SEEDED_CASES[case]
    boxes, sigt_mat, sigs_mat = seeded_inputs(kw["seed"])
    o = _oracle(kw, layout)
    G = kw["groups"]
    nr = kw["zset"][0] * kw["zset"][1] * kw["zset"][2]
    for r in range(nr):  # one group set: PR chunk id == R chunk id
        want = expected_sigt(expected_mix_seeded(kw, r, boxes), sigt_mat)
        got = o.get("sigt_zonal", r)
        # Field_SigmaTZonal <Group, Zone>: a one-element "direction" axis in the flux-layout helper
        got = canonical(got, layout, 1, G, want.size)[0]
        for g in range(G):
            assert np.array_equal(got[g], want), (r, g)
    sigs = o.get("sigs").reshape(3, 2, G, G)  # oracle layout [mat][n][g][gp]
    want = np.zeros((3, 2, G, G))
    for m in range(3):
        want[m, 0][np.arange(G), np.arange(G)] = sigs_mat[m]
    assert np.array_equal(sigs, want)
    # a sigma_s-for-sigma_t mix-up, or a wrong material order, would differ here
    assert not np.allclose(expected_sigt(expected_mix_seeded(kw, 0, boxes), sigs_mat), got[0])


def test_product_generator_against_the_recomputed_assembly():
    """The host mirror's generator (device-free problem) against the same
    recomputation: fractions, sigt_zonal and sigs."""
    kw = SEEDED_CASES[0]
    boxes, sigt_mat, sigs_mat = seeded_inputs(kw["seed"])
    hp = ka.HostProblem(ka.input_vars(quad=(4, 4), legendre=1, dset=8, gset=1, geometry=1, layout="DGZ", **kw))
    try:
        G = kw["groups"]
        for r in range(hp.num_r):
            frac = expected_mix_seeded(kw, r, boxes)
            first, num, zone, mat, fr, _ = mix_tables(frac)
            assert np.array_equal(hp.get("mixelem_to_fraction", r), fr)
            assert np.array_equal(hp.get("mixelem_to_material", r), mat)
            assert np.array_equal(hp.get("zone_to_mixelem", r), first)
            want = expected_sigt(frac, sigt_mat)
            got = hp.get("sigt_zonal", r).reshape(G, -1)
            for g in range(G):
                assert np.array_equal(got[g], want)
        s = hp.get("data/sigs").reshape(3, 2, G, G)
        for m in range(3):
            assert np.array_equal(np.diag(s[m, 0]), np.full(G, sigs_mat[m]))
            assert np.count_nonzero(s[m]) == G
    finally:
        hp.close()


# Kobayashi-3i (Space.cpp:142-177): region table over the first octant,
# with |x|, |y|, |z|; each row (material, xlo, xhi, ylo, yhi, zlo, zhi),
# closed intervals, first match wins.
KOBAYASHI = [
    (0, 0, 10, 0, 10, 0, 10),
    (1, 0, 10, 0, 60, 0, 10),
    (1, 0, 40, 50, 60, 0, 10),
    (1, 30, 40, 50, 60, 0, 40),
    (1, 30, 40, 50, np.inf, 30, 40),
]


def kobayashi_fractions(nx, ny, nz, ns=4):
    """[zone, material] over the whole mesh (one zone set), evaluating the
    region table on the full sub-sample grid."""
    dx, dy, dz = 120.0 / nx, 200.0 / ny, 120.0 / nz
    xs = axis_samples(-60.0 + dx * 0, dx, nx, ns).ravel()
    ys = axis_samples(-100.0 + dy * 0, dy, ny, ns).ravel()
    zs = axis_samples(-60.0 + dz * 0, dz, nz, ns).ravel()
    Z, Y, X = np.meshgrid(np.abs(zs), np.abs(ys), np.abs(xs), indexing="ij")
    mat = np.full(X.shape, 2, dtype=np.int64)
    done = np.zeros(X.shape, dtype=bool)
    for m, x0, x1, y0, y1, z0, z1 in KOBAYASHI:
        hit = ~done & (X >= x0) & (X <= x1) & (Y >= y0) & (Y <= y1) & (Z >= z0) & (Z <= z1)
        mat[hit] = m
        done |= hit
    # [k, sk, j, sj, i, si] -> counts per zone and material
    mat = mat.reshape(nz, ns, ny, ns, nx, ns)
    cnt = np.stack([np.sum(mat == m, axis=(1, 3, 5)) for m in range(3)], axis=-1)
    return cnt.reshape(-1, 3) / float(ns ** 3)


@pytest.mark.parametrize("zones", [(12, 20, 12), (9, 14, 7)])
def test_kobayashi_assembly(zones):
    sigt = (0.1, 0.0001, 0.1)
    o = po.Oracle(po.make_input(zones=zones, groups=2, quad=None, directions=8, legendre=0, geometry=0, sigt=sigt))
    frac = kobayashi_fractions(*zones)
    first, num, zone, mat, fr, _ = mix_tables(frac)
    assert np.array_equal(o.get("mixelem_to_material", 0), mat)
    assert np.array_equal(o.get("mixelem_to_fraction", 0), fr)
    assert np.array_equal(o.get("zone_to_num_mixelem", 0), num)
    got = o.get("sigt_zonal", 0).reshape(2, -1)
    want = expected_sigt(frac, np.array(sigt))
    assert np.array_equal(got[0], want) and np.array_equal(got[1], want)
    if zones == (12, 20, 12):
        # zone faces on every region boundary, no sample on one: the sampled
        # volumes are the exact region volumes, 8 octants of
        # 10^3 (material 0) and 5000 + 3000 + 3000 + 4000 (material 1)
        vol = o.get("volume", 0)[0]
        v = frac.sum(axis=0) * vol
        assert np.allclose(v, [8000.0, 120000.0, 120.0 * 200.0 * 120.0 - 128000.0], rtol=1e-13, atol=0)
