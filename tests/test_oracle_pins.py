"""Pins of the CPU oracle (no GPU).

The reference ships no tests, golden values or fixtures for this path, and it
cannot be built here (its RAJA/camp/BLT submodules are empty), so every oracle
function is pinned by something it does not share code with: numpy/scipy
routines, closed forms, mathematical invariants, and a brute-force global
linear solve of the diamond-difference equations.
"""
import math

import numpy as np
import pytest
from scipy.special import lpmv

import pyoracle as po
from bruteforce import (blockwise_sweep_numpy, global_sweep_numpy, lplustimes_numpy, ltimes_numpy, mix_maps,
                        numpy_iteration, population_fsum)
from helpers import LAYOUT_AXES, canonical, rel_normwise

LAYOUTS = list(LAYOUT_AXES)


def mk(kind="ref", omp=True, **kw):
    base = dict(zones=(4, 4, 4), groups=4, quad=(4, 4), legendre=2, layout="DGZ", dset=8, gset=1,
                zset=(1, 1, 1), geometry=1, seed=7)
    base.update(kw)
    return po.Oracle(po.make_input(**base), kind=kind, omp=omp)


def all_directions(o):
    """(signed cosines, w, ell, ell_plus) over every direction set, in order."""
    rows = []
    for q in range(o.inp.num_dirsets):
        xc, yc, zc, w = o.get("xcos", q), o.get("ycos", q), o.get("zcos", q), o.get("w", q)
        sx, sy, sz = o.get("id", q), o.get("jd", q), o.get("kd", q)
        for d in range(xc.size):
            rows.append((sx[d] * xc[d], sy[d] * yc[d], sz[d] * zc[d], w[d]))
    return np.array(rows)


# --------------------------------------------------------------------------
# quadrature and spherical harmonics
# --------------------------------------------------------------------------
@pytest.mark.parametrize("quad", [(4, 4), (8, 12), (8, 16), (12, 8)])
def test_gauss_legendre_vs_numpy(quad):
    np_, na = quad
    o = mk(quad=quad, zones=(2, 2, 2), groups=1, legendre=0)
    dirs = all_directions(o)
    polar = np.unique(np.round(dirs[:, 2], 15))
    x, w = np.polynomial.legendre.leggauss(np_)
    assert np.allclose(np.sort(polar), np.sort(x), rtol=0, atol=2e-15)
    dangle = 2 * np.pi / na
    for mu, wt in zip(x, w):
        sel = np.abs(dirs[:, 2] - mu) < 1e-12
        assert sel.sum() == na
        assert np.allclose(dirs[sel, 3], wt * dangle, rtol=2e-14, atol=0)


def _ynm_scipy(n, m, mu, eta, xi):
    phi = math.atan2(eta, mu) % (2 * math.pi)
    am = abs(m)
    a = math.sqrt((2 * n + 1) * math.factorial(n - am) / (2 * (1 + (m == 0)) * math.pi * math.factorial(n + am)))
    p = lpmv(am, n, xi)  # includes the Condon-Shortley phase
    return a * p * (math.cos(m * phi) if m >= 0 else math.sin(am * phi))


@pytest.mark.parametrize("quad", [(8, 12), (4, 8)])
def test_spherical_harmonics_vs_scipy(quad):
    L = 4
    o = mk(quad=quad, legendre=L, zones=(2, 2, 2), groups=1)
    M = (L + 1) ** 2
    dirs = all_directions(o)
    elp = np.concatenate([o.get("ell_plus", q).reshape(-1, M) for q in range(8)])
    worst = 0.0
    for d, (mu, eta, xi, _) in enumerate(dirs):
        nm = 0
        for n in range(L + 1):
            for m in range(-n, n + 1):
                y = _ynm_scipy(n, m, mu, eta, xi)
                worst = max(worst, abs(elp[d, nm] / math.sqrt(4 * math.pi) - y))
                nm += 1
    assert worst < 2e-14


@pytest.mark.parametrize("quad,dirs", [((8, 12), None), ((4, 4), None), (None, 96), (None, 16)])
def test_ell_closed_forms(quad, dirs):
    o = mk(quad=quad, directions=dirs, legendre=3, zones=(2, 2, 2), groups=1)
    M = 16
    ws, e0, ep0 = [], [], []
    for q in range(8):
        w = o.get("w", q)
        ell = o.get("ell", q).reshape(-1, M)
        elp = o.get("ell_plus", q).reshape(-1, M)
        ws.append(w)
        e0.append(ell[:, 0] - w / (4 * np.pi))
        ep0.append(elp[:, 0] - 1.0)
    assert abs(math.fsum(np.concatenate(ws)) - 4 * np.pi) < 1e-13
    assert np.max(np.abs(np.concatenate(e0))) < 1e-17
    assert np.max(np.abs(np.concatenate(ep0))) < 4e-16


@pytest.mark.parametrize("quad", [(8, 12), (8, 16)])
def test_orthonormality_gauss_legendre(quad):
    """For the GL product set with P polar and A azimuthal points, P4 moments
    are integrated exactly: L L+ = I."""
    o = mk(quad=quad, legendre=4, zones=(2, 2, 2), groups=1)
    M = 25
    ell = np.concatenate([o.get("ell", q).reshape(-1, M) for q in range(8)])
    elp = np.concatenate([o.get("ell_plus", q).reshape(-1, M) for q in range(8)])
    assert np.max(np.abs(ell.T @ elp - np.eye(M))) < 1e-13


def test_orthonormality_fails_for_dummy_s2():
    """The dummy S2 set repeats one direction per octant; it cannot integrate
    P4 moments, and the check must see that."""
    o = mk(quad=None, directions=96, legendre=4, zones=(2, 2, 2), groups=1)
    M = 25
    ell = np.concatenate([o.get("ell", q).reshape(-1, M) for q in range(8)])
    elp = np.concatenate([o.get("ell_plus", q).reshape(-1, M) for q in range(8)])
    assert np.max(np.abs(ell.T @ elp - np.eye(M))) > 0.1


def test_direction_sets_are_octants():
    """dset q has id=+1 iff q&1, jd=+1 iff q&2, kd=+1 iff q&4 for both sets."""
    for quad, dirs in (((8, 12), None), (None, 96)):
        o = mk(quad=quad, directions=dirs, zones=(2, 2, 2), groups=1)
        for q in range(8):
            assert set(o.get("id", q)) == {1 if q & 1 else -1}
            assert set(o.get("jd", q)) == {1 if q & 2 else -1}
            assert set(o.get("kd", q)) == {1 if q & 4 else -1}


# --------------------------------------------------------------------------
# moment kernels
# --------------------------------------------------------------------------
@pytest.mark.parametrize("kind", ["ref", "fma"])
@pytest.mark.parametrize("layout", LAYOUTS)
def test_ltimes_lplustimes_vs_einsum(layout, kind):
    o = mk(kind=kind, layout=layout, gset=2, zset=(2, 1, 1), groups=4)
    rng = np.random.default_rng(3)
    M = 9
    for s in range(o.num_sdoms):
        o.set("psi", s, rng.standard_normal(o.get("psi", s).size))
    for c in range(o.num_pr):
        o.set("phi_out", c, rng.standard_normal(o.get("phi_out", c).size))
    o.kconst("phi", 0.0)
    o.kconst("rhs", 0.0)
    o.ltimes()
    o.lplustimes()
    ref_phi = ltimes_numpy(o, layout, M)
    ref_rhs = lplustimes_numpy(o, layout, M)
    S0 = o.sdom(0)
    nz = S0["ni"] * S0["nj"] * S0["nk"]
    for c in range(o.num_pr):
        got = canonical(o.get("phi", c), layout, M, S0["ng"], nz)
        assert rel_normwise(got, ref_phi[c]) < 1e-14
    for s in range(o.num_sdoms):
        got = canonical(o.get("rhs", s), layout, S0["nd"], S0["ng"], nz)
        assert rel_normwise(got, ref_rhs[s]) < 1e-14


@pytest.mark.parametrize("layout", ["DGZ", "ZGD", "GDZ"])
def test_scattering_vs_dense_definition(layout):
    """phi_out(nm,g,z) = sum_gp [sum_mix frac*sigs(mat,n,g,gp)] phi(nm,gp,z)
    evaluated densely in numpy, with a coupled, anisotropic sigs."""
    o = mk(layout=layout, gset=2, groups=4, legendre=1, zset=(1, 2, 1))
    rng = np.random.default_rng(4)
    L1, G, M = 2, 4, 4
    sigs = rng.uniform(0, 0.2, (3, L1, G, G))
    o.set("sigs", 0, sigs.ravel())
    for c in range(o.num_pr):
        o.set("phi", c, rng.standard_normal(o.get("phi", c).size))
    o.kconst("phi_out", 0.0)
    o.scattering()
    m2l = np.array([0, 1, 1, 1])
    S0 = o.sdom(0)
    nz = S0["ni"] * S0["nj"] * S0["nk"]
    ng = S0["ng"]
    # group the PR chunks by zone set: phi over all global groups
    chunks = {}
    for s in range(o.num_sdoms):
        S = o.sdom(s)
        chunks.setdefault(S["r"], {})[S["gset"]] = S["pr"]
    for r, by_g in chunks.items():
        z2m, z2n, mat, frac = mix_maps(o, r)
        sz = np.zeros((nz, L1, G, G))
        for z in range(nz):
            for k in range(z2n[z]):
                sz[z] += frac[z2m[z] + k] * sigs[mat[z2m[z] + k]]
        phig = np.concatenate([canonical(o.get("phi", by_g[gs]), layout, M, ng, nz) for gs in sorted(by_g)], axis=1)
        want = np.einsum("zmgh,mhz->mgz", sz[:, m2l], phig)
        for gs in sorted(by_g):
            got = canonical(o.get("phi_out", by_g[gs]), layout, M, ng, nz)
            assert rel_normwise(got, want[:, gs * ng:(gs + 1) * ng]) < 1e-14


@pytest.mark.parametrize("geometry", [0, 1])
def test_source_closed_form(geometry):
    o = mk(geometry=geometry, zones=(8, 8, 8), groups=3, zset=(2, 1, 1), legendre=1)
    o.kconst("phi_out", 0.0)
    o.source()
    S0 = o.sdom(0)
    nz = S0["ni"] * S0["nj"] * S0["nk"]
    for s in range(o.num_sdoms):
        S = o.sdom(s)
        if S["dset"] != 0:
            continue
        _, _, mat, frac = mix_maps(o, S["r"])
        want = math.fsum(frac[mat == 0])
        phi_out = canonical(o.get("phi_out", S["pr"]), "DGZ", 4, 3, nz)
        for g in range(3):
            assert math.fsum(phi_out[0, g]) == pytest.approx(want, rel=1e-15, abs=0)
        assert not np.any(phi_out[1:])


# --------------------------------------------------------------------------
# sweep
# --------------------------------------------------------------------------
@pytest.mark.parametrize("quad,dirs", [((4, 4), None), (None, 16)])
@pytest.mark.parametrize("layout", LAYOUTS)
def test_sweep_vs_global_linear_solve(layout, quad, dirs):
    zones = (4, 2, 4)
    o = mk(layout=layout, quad=quad, directions=dirs, zones=zones, zset=(2, 1, 2), groups=2, gset=2)
    rng = np.random.default_rng(5)
    for s in range(o.num_sdoms):
        o.set("rhs", s, rng.uniform(0.5, 2.0, o.get("rhs", s).size))
    o.sweep_solver()
    want = global_sweep_numpy(o, layout, *zones)
    S0 = o.sdom(0)
    nz = S0["ni"] * S0["nj"] * S0["nk"]
    for s in range(o.num_sdoms):
        got = canonical(o.get("psi", s), layout, S0["nd"], S0["ng"], nz)
        assert rel_normwise(got, want[s]) < 1e-12


@pytest.mark.parametrize("layout", ["DGZ", "GZD", "ZDG"])
def test_block_jacobi_vs_blockwise_numpy(layout):
    """--pmethod bj on one rank: each subdomain swept alone with zero
    incoming faces (the communicator's fresh old planes), against a dense
    per-block diamond-difference solve with vacuum inflow."""
    zones = (4, 6, 4)
    o = mk(layout=layout, zones=zones, zset=(2, 2, 1), groups=2, gset=2)
    rng = np.random.default_rng(8)
    for s in range(o.num_sdoms):
        o.set("rhs", s, rng.uniform(0.5, 2.0, o.get("rhs", s).size))
    o.sweep_solver(block_jacobi=True)
    want = blockwise_sweep_numpy(o, layout)
    S0 = o.sdom(0)
    nz = S0["ni"] * S0["nj"] * S0["nk"]
    for s in range(o.num_sdoms):
        got = canonical(o.get("psi", s), layout, S0["nd"], S0["ng"], nz)
        assert rel_normwise(got, want[s]) < 1e-12
    # the coupled sweep differs (interior faces carry flux)
    o.sweep_solver()
    diff = max(rel_normwise(canonical(o.get("psi", s), layout, S0["nd"], S0["ng"], nz), want[s])
               for s in range(o.num_sdoms))
    assert diff > 1e-3


def test_block_jacobi_one_zone_set_is_the_sweep():
    """With one zone set there are no interior faces: Block Jacobi iterations
    equal the sweep iterations bit for bit."""
    a = mk(zones=(4, 4, 4), zset=(1, 1, 1), pmethod="bj", omp=False)
    b = mk(zones=(4, 4, 4), zset=(1, 1, 1), omp=False)
    a.init_psi()
    b.init_psi()
    for _ in range(3):
        pa, pb = a.iterate(), b.iterate()
        assert pa == pb
    for s in range(a.num_sdoms):
        assert np.array_equal(a.get("psi", s), b.get("psi", s))


def test_block_jacobi_iteration_population_lower():
    """With several zone sets Block Jacobi drops the inflow across zone-set
    faces, so every iteration's particle count is below the sweep's (positive
    source, psi from zero)."""
    a = mk(zones=(8, 8, 8), zset=(2, 2, 2), pmethod="bj")
    b = mk(zones=(8, 8, 8), zset=(2, 2, 2))
    a.init_psi()
    b.init_psi()
    for _ in range(3):
        pa, pb = a.iterate()[1], b.iterate()[1]
        assert 0.0 < pa < pb


def test_sweep_single_zone_closed_form():
    o = mk(zones=(1, 1, 1), groups=2, quad=(4, 4), legendre=0)
    rng = np.random.default_rng(6)
    for s in range(o.num_sdoms):
        o.set("rhs", s, rng.uniform(1, 2, o.get("rhs", s).size))
    o.sweep_solver()
    st = o.get("sigt_zonal", 0)
    dx, dy, dz = o.get("dx")[0], o.get("dy")[0], o.get("dz")[0]
    for s in range(o.num_sdoms):
        q = o.sdom(s)["dset"]
        xc, yc, zc = o.get("xcos", q), o.get("ycos", q), o.get("zcos", q)
        rhs = o.get("rhs", s).reshape(2, 2)  # [d][g] for one zone
        psi = o.get("psi", s).reshape(2, 2)
        for d in range(2):
            for g in range(2):
                den = 2 * xc[d] / dx + 2 * yc[d] / dy + 2 * zc[d] / dz + st[g]
                assert psi[d, g] == pytest.approx(rhs[d, g] / den, rel=2e-16)
        for name in ("i_plane", "j_plane", "k_plane"):
            assert np.array_equal(o.get(name, s), 2.0 * o.get("psi", s))


def test_sweep_zero_source_and_linearity():
    o = mk(zones=(4, 4, 4), zset=(2, 2, 1), groups=2)
    rng = np.random.default_rng(8)
    o.kconst("rhs", 0.0)
    o.sweep_solver()
    for s in range(o.num_sdoms):
        assert not np.any(o.get("psi", s))
    q1 = [rng.standard_normal(o.get("rhs", s).size) for s in range(o.num_sdoms)]
    q2 = [rng.standard_normal(o.get("rhs", s).size) for s in range(o.num_sdoms)]
    res = []
    for qs in (q1, q2, [3.0 * a - 0.5 * b for a, b in zip(q1, q2)]):
        for s in range(o.num_sdoms):
            o.set("rhs", s, qs[s])
        o.sweep_solver()
        res.append([o.get("psi", s) for s in range(o.num_sdoms)])
    for s in range(o.num_sdoms):
        assert rel_normwise(res[2][s], 3.0 * res[0][s] - 0.5 * res[1][s]) < 1e-13


def test_sweep_mirror_symmetry_bitwise():
    """Kobayashi geometry is symmetric under x -> -x; with the dummy S2 set the
    octants differing in id see mirrored problems, so psi(q, i) equals
    psi(q^1, ni-1-i) bit for bit after an iteration from psi = 0."""
    o = mk(geometry=0, quad=None, directions=8, zones=(6, 4, 4), groups=2, legendre=1)
    o.init_psi()
    o.iterate()
    for q in range(8):
        a = o.get("psi", q).reshape(1, 2, 4, 4, 6)   # DGZ, one direction per set
        b = o.get("psi", q ^ 1).reshape(1, 2, 4, 4, 6)
        assert np.array_equal(a, b[..., ::-1])


def test_population_vs_fsum():
    o = mk(zones=(4, 4, 4), zset=(1, 2, 1), gset=2, groups=4)
    rng = np.random.default_rng(9)
    for s in range(o.num_sdoms):
        o.set("psi", s, rng.uniform(0, 5, o.get("psi", s).size))
    lit, comp = o.population()
    exact = population_fsum(o, "DGZ")
    assert abs(comp - exact) <= 2e-16 * abs(exact)
    assert abs(lit - exact) <= 1e-13 * abs(exact)


# --------------------------------------------------------------------------
# whole iterations
# --------------------------------------------------------------------------
@pytest.mark.parametrize("layout", ["DGZ", "GZD", "ZDG"])
def test_full_iterations_vs_numpy(layout):
    kw = dict(layout=layout, zones=(4, 4, 4), zset=(2, 1, 2), gset=2, groups=4, quad=(4, 4), legendre=2)
    o = mk(**kw)      # the oracle under test
    n = mk(**kw)      # scratch oracle holding the numpy iteration's state
    o.init_psi()
    n.init_psi()
    M, m2l = 9, np.array([0, 1, 1, 1, 2, 2, 2, 2, 2])
    S0 = o.sdom(0)
    nz = S0["ni"] * S0["nj"] * S0["nk"]
    for it in range(3):
        _, comp = o.iterate()
        phi_np, psi_np = numpy_iteration(n, layout, M, m2l)
        for s, v in psi_np.items():
            n.set("psi", s, np.transpose(v, [("agz").index(ch) for ch in LAYOUT_AXES[layout]]).ravel())
        assert comp == pytest.approx(population_fsum(n, layout), rel=1e-12)
        for c in range(o.num_pr):
            got = canonical(o.get("phi", c), layout, M, S0["ng"], nz)
            assert rel_normwise(got, phi_np[c]) < 1e-12


def test_population_increases_from_zero():
    o = mk(zones=(8, 8, 8), groups=4, quad=(4, 8), legendre=2)
    o.init_psi()
    pops = [o.iterate()[1] for _ in range(5)]
    assert all(b > a for a, b in zip(pops, pops[1:]))


def test_decomposition_invariance():
    """Zone-set, group-set and direction-set decompositions change only the
    summation order (LTimes over direction sets) and the sweep schedule."""
    base = dict(zones=(8, 8, 8), groups=4, quad=(8, 4), legendre=2)
    runs = [dict(), dict(zset=(2, 2, 2)), dict(gset=2), dict(dset=16, zset=(1, 2, 1))]
    out = []
    for extra in runs:
        o = mk(**dict(base, **extra))
        o.init_psi()
        out.append([o.iterate()[1] for _ in range(3)])
    for other in out[1:]:
        for a, b in zip(out[0], other):
            assert a == pytest.approx(b, rel=1e-13)


def test_nesting_invariance_bitwise():
    ref = None
    for layout in LAYOUTS:
        o = mk(layout=layout, zones=(4, 4, 4), zset=(2, 1, 1), groups=4)
        o.init_psi()
        for _ in range(2):
            o.iterate()
        S0 = o.sdom(0)
        psi = [canonical(o.get("psi", s), layout, S0["nd"], S0["ng"], 32) for s in range(o.num_sdoms)]
        if ref is None:
            ref = psi
        else:
            for a, b in zip(psi, ref):
                assert np.array_equal(a, b)


def test_omp_and_sequential_agree():
    """OpenMP loop orders change nothing but the population's reduction order."""
    kw = dict(zones=(8, 8, 8), groups=4, zset=(2, 1, 1), layout="GZD")
    a, b = mk(omp=False, **kw), mk(omp=True, **kw)
    a.init_psi()
    b.init_psi()
    for _ in range(2):
        la, ca = a.iterate()
        lb, cb = b.iterate()
        assert ca == cb
    for c in range(a.num_pr):
        assert np.array_equal(a.get("phi", c), b.get("phi", c))
