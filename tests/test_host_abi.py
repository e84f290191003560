"""CPU-only checks of the product: the C-ABI library loads and exports every
function include/kripke_amd.h declares, error behaviour without a device,
and the host mirror's generateProblem reproduces the oracle's setup fields
bit for bit (no GPU compute is called here)."""
import ctypes
import os
import re

import numpy as np
import pytest

import kripke_amd as ka
import pyoracle as po
from kripke_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "kripke_amd.h")


def header_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(kr_[a-z_0-9]+)\s*\(", src, flags=re.M)))


def test_header_declares_expected_entry_points():
    names = header_functions()
    for must in ("kr_ltimes", "kr_lplustimes", "kr_scattering", "kr_source", "kr_sweep_subdomain",
                 "kr_sweep_solver", "kr_population", "kr_kconst", "kr_kcopy", "kr_problem_create"):
        assert must in names


def test_library_exports_every_declared_symbol():
    L = ka.lib()
    for name in header_functions():
        assert hasattr(L, name), name
    bound = {n for n, _, _ in capi.SIGNATURES}
    assert bound == set(header_functions())


def test_no_device_errors_cleanly():
    if ka.device_count() > 0:
        pytest.skip("a device is visible")
    with pytest.raises(ka.KripkeError) as e:
        ka.Context(0)
    assert e.value.code == 5  # KR_ERR_NO_DEVICE
    # entry points reject null handles instead of crashing
    assert ka.lib().kr_ltimes(None, 1) == 1
    assert ka.lib().kr_population(None, None) == 1


@pytest.mark.parametrize("bad", [
    dict(zones=(10, 8, 8), zset=(3, 1, 1)),        # zones not divisible by zone sets (Space.cpp:49)
    dict(groups=6, gset=4),                        # groups not divisible by group sets
    dict(quad=(6, 4)),                             # polar not a multiple of 4
    dict(legendre=-1),
    dict(layout=7),
])
def test_invalid_inputs_rejected(bad):
    kw = dict(zones=(8, 8, 8), groups=8, quad=(4, 4), legendre=2)
    kw.update(bad)
    with pytest.raises(ka.KripkeError) as e:
        ka.HostProblem(ka.input_vars(**kw))
    assert e.value.code == 1


GEN_FIELDS = [
    ("ell", "ell", "q"), ("ell_plus", "ell_plus", "q"), ("quadrature/xcos", "xcos", "q"),
    ("quadrature/ycos", "ycos", "q"), ("quadrature/zcos", "zcos", "q"), ("quadrature/w", "w", "q"),
    ("quadrature/id", "id", "q"), ("quadrature/jd", "jd", "q"), ("quadrature/kd", "kd", "q"),
    ("sigt_zonal", "sigt_zonal", "pr"), ("volume", "volume", "r"), ("mixelem_to_zone", "mixelem_to_zone", "r"),
    ("mixelem_to_material", "mixelem_to_material", "r"), ("mixelem_to_fraction", "mixelem_to_fraction", "r"),
    ("zone_to_mixelem", "zone_to_mixelem", "r"), ("zone_to_num_mixelem", "zone_to_num_mixelem", "r"),
]


@pytest.mark.parametrize("cfg", [
    dict(layout="DGZ"), dict(layout="GZD", zset=(2, 1, 2), gset=2), dict(layout="ZDG", seed=99, quad=(8, 12)),
    dict(layout="DZG", geometry=0, quad=None, directions=32), dict(layout="ZGD", legendre=4, zones=(12, 8, 4)),
])
def test_generator_matches_oracle_bitwise(cfg):
    kw = dict(zones=(8, 8, 8), groups=4, quad=(4, 4), legendre=2, dset=8, gset=1, zset=(1, 1, 1), geometry=1,
              seed=20261202)
    kw.update(cfg)
    hp = ka.HostProblem(ka.input_vars(**kw))  # host-only: no device needed
    o = po.Oracle(po.make_input(**kw))
    counts = {"q": 8, "pr": o.num_pr, "r": hp.num_r}
    for pname, oname, space in GEN_FIELDS:
        for c in range(counts[space]):
            a, b = hp.get(pname, c), o.get(oname, c)
            assert np.array_equal(a, b), (pname, c)
    # data/sigs: oracle keeps [mat][n][g][gp]; the field is in the nesting's order
    sig = hp.get("data/sigs")
    L1, G = kw["legendre"] + 1, kw["groups"]
    n_first = kw["layout"] in ("DGZ", "DZG", "ZDG")
    s = sig.reshape(3, L1, G, G) if n_first else np.transpose(sig.reshape(3, G, G, L1), (0, 3, 1, 2))
    assert np.array_equal(np.ascontiguousarray(s).ravel(), o.get("sigs"))
    dx = np.concatenate([hp.get("dx", c) for c in range(kw["zset"][0])])
    assert np.array_equal(dx, o.get("dx"))


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


def test_splitmix64_known_answer():
    # published first output of SplitMix64 from state 0
    assert _splitmix64(0, 1)[0] == 0xE220A8397B1DCDAF


@pytest.mark.parametrize("seed", [20261202, 1, 12345])
def test_seeded_materials_follow_the_stated_generator(seed):
    """Cross sections and box geometry from an independent Python SplitMix64."""
    u = [(x >> 11) * 2.0 ** -53 for x in _splitmix64(seed, 18)]
    boxes = []
    for b in range(2):
        lo, hi = [], []
        for ax in range(3):
            side = 100.0 * (0.1 + 0.4 * u[b * 6 + 2 * ax])
            l0 = (100.0 - side) * u[b * 6 + 2 * ax + 1]
            lo.append(l0)
            hi.append(l0 + side)
        boxes.append((lo, hi))
    sigt = [10.0 ** (-4.0 + 4.0 * u[12 + m]) for m in range(3)]
    sigs = [(0.3 + 0.6 * u[15 + m]) * sigt[m] for m in range(3)]
    o = po.Oracle(po.make_input(zones=(10, 10, 10), groups=1, quad=(4, 4), legendre=0, seed=seed))
    st, ss = o.materials()
    assert np.allclose(st, sigt, rtol=1e-15, atol=0)
    assert np.allclose(ss, sigs, rtol=1e-15, atol=0)
    # material-0 volume equals the fraction-weighted count of sub-samples in box 0
    z2m, z2n = o.get("zone_to_mixelem", 0), o.get("zone_to_num_mixelem", 0)
    frac, mat = o.get("mixelem_to_fraction", 0), o.get("mixelem_to_material", 0)
    per_zone = np.zeros(1000)
    np.add.at(per_zone, o.get("mixelem_to_zone", 0).astype(int), frac)
    assert np.all(per_zone == 1.0)
    assert np.array_equal(z2m[1:], np.cumsum(z2n)[:-1])
    lo, hi = boxes[1]
    xs = (np.arange(10)[:, None] * 10.0 + 10.0 * (np.arange(4)[None, :] + 1) / 5).ravel()
    inside = [((xs >= lo[a]) & (xs <= hi[a])).astype(float) for a in range(3)]
    n0 = np.einsum("i,j,k->", inside[2], inside[1], inside[0]) / 64.0
    assert math_fsum(frac[mat == 0]) == pytest.approx(n0, rel=1e-12)


def math_fsum(x):
    import math
    return math.fsum(x)


def test_kobayashi_material_volumes():
    """The reference default geometry (Space.cpp:142-177): central 20^3 box
    is material 0, so its volume is 8000 cm^3 with zone faces on the box."""
    o = po.Oracle(po.make_input(zones=(12, 20, 12), groups=1, quad=None, directions=8, legendre=0, geometry=0))
    frac, mat = o.get("mixelem_to_fraction", 0), o.get("mixelem_to_material", 0)
    vol = o.get("volume", 0)[0]
    assert math_fsum(frac[mat == 0]) * vol == pytest.approx(8000.0, rel=1e-12)
