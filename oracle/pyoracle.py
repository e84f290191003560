"""ORACLE — test infrastructure only.

ctypes wrapper over the CPU oracle (oracle/kripke_oracle.cpp).  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg import this module;
the product (kripke_amd/) never does.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(HERE, "_build")

LAYOUTS = ["DGZ", "DZG", "GDZ", "GZD", "ZDG", "ZGD"]


class KoInput(ctypes.Structure):
    """Mirror of ko_input (oracle/kripke_oracle.h)."""

    _fields_ = [
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
        ("num_directions", ctypes.c_int32),
        ("quad_num_polar", ctypes.c_int32), ("quad_num_azimuthal", ctypes.c_int32),
        ("num_groups", ctypes.c_int32), ("legendre_order", ctypes.c_int32),
        ("layout", ctypes.c_int32),
        ("num_dirsets", ctypes.c_int32), ("num_groupsets", ctypes.c_int32),
        ("num_zonesets", ctypes.c_int32 * 3),
        ("num_material_subsamples", ctypes.c_int32),
        ("geometry", ctypes.c_int32),
        ("seed", ctypes.c_uint64),
        ("sigt", ctypes.c_double * 3), ("sigs", ctypes.c_double * 3),
        ("source_strength", ctypes.c_double),
        ("parallel_method", ctypes.c_int32),
    ]


def make_input(zones=(16, 16, 16), groups=32, quad=(8, 12), directions=None, legendre=4,
               layout="DGZ", dset=8, gset=1, zset=(1, 1, 1), geometry=1, seed=20261202,
               subsamples=4, sigt=(0.1, 0.0001, 0.1), sigs=(0.05, 0.00005, 0.05),
               source_strength=1.0, pmethod="sweep"):
    """Problem description with the reference's InputVariables meaning
    (src/Kripke/InputVariables.cpp:17-46).  quad=(polar, azimuthal) selects
    the Gauss-Legendre product set, quad=None the dummy S2 set with
    `directions` points; pmethod is --pmethod (sweep | bj, kripke.cpp:329-340)."""
    ki = KoInput()
    ki.nx, ki.ny, ki.nz = zones
    if quad is None:
        ki.num_directions = directions
        ki.quad_num_polar = ki.quad_num_azimuthal = 0
    else:
        ki.quad_num_polar, ki.quad_num_azimuthal = quad
        ki.num_directions = quad[0] * quad[1]
    ki.num_groups = groups
    ki.legendre_order = legendre
    ki.layout = LAYOUTS.index(layout) if isinstance(layout, str) else int(layout)
    ki.num_dirsets = dset
    ki.num_groupsets = gset
    for a in range(3):
        ki.num_zonesets[a] = zset[a]
    ki.num_material_subsamples = subsamples
    ki.geometry = geometry
    ki.seed = seed
    for m in range(3):
        ki.sigt[m] = sigt[m]
        ki.sigs[m] = sigs[m]
    ki.source_strength = source_strength
    ki.parallel_method = {"sweep": 0, "bj": 1}[pmethod]
    return ki


_libs = {}


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib(kind="ref"):
    """kind = 'ref' (reference arithmetic) or 'fma' (HIP accumulation order)."""
    if kind in _libs:
        return _libs[kind]
    path = os.path.join(BUILD, "libkripke_oracle_%s.so" % kind)
    if not os.path.exists(path):
        build()
    L = ctypes.CDLL(path)
    P = ctypes.c_void_p
    L.ko_create.restype = P
    L.ko_create.argtypes = [ctypes.POINTER(KoInput)]
    L.ko_destroy.argtypes = [P]
    L.ko_error.restype = ctypes.c_char_p
    L.ko_num_sdoms.argtypes = [P]
    L.ko_num_pr.argtypes = [P]
    L.ko_sdom_info.argtypes = [P, ctypes.c_int32, ctypes.POINTER(ctypes.c_int64)]
    L.ko_materials.argtypes = [P, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    L.ko_get.restype = ctypes.c_int64
    L.ko_get.argtypes = [P, ctypes.c_char_p, ctypes.c_int32, ctypes.POINTER(ctypes.c_double)]
    L.ko_set.restype = ctypes.c_int64
    L.ko_set.argtypes = [P, ctypes.c_char_p, ctypes.c_int32, ctypes.POINTER(ctypes.c_double)]
    L.ko_kconst_all.argtypes = [P, ctypes.c_char_p, ctypes.c_double]
    for fn in ("ko_ltimes", "ko_lplustimes", "ko_scattering", "ko_source", "ko_sweep_solver"):
        getattr(L, fn).argtypes = [P, ctypes.c_int]
    L.ko_sweep_subdomain.argtypes = [P, ctypes.c_int32, ctypes.c_int]
    L.ko_sweep_solver_bj.argtypes = [P, ctypes.c_int]
    dp = ctypes.POINTER(ctypes.c_double)
    L.ko_population.argtypes = [P, ctypes.c_int, dp, dp]
    L.ko_iterate.argtypes = [P, ctypes.c_int, dp, dp]
    L.ko_init_psi.argtypes = [P]
    L.ko_timer.restype = ctypes.c_double
    L.ko_timer.argtypes = [P, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int64)]
    L.ko_timer_reset.argtypes = [P]
    L.ko_omp_threads.restype = ctypes.c_int32
    _libs[kind] = L
    return L


SDOM_KEYS = ["gset", "dset", "zx", "zy", "zz", "nd", "ng", "ni", "nj", "nk", "dlo", "glo", "ilo",
             "jlo", "klo", "pr", "r", "up0", "up1", "up2", "down0", "down1", "down2", "id", "jd", "kd"]


class Oracle:
    """One generated problem held by the CPU oracle."""

    def __init__(self, inp, kind="ref", omp=False):
        self.L = lib(kind)
        self.inp = inp
        self.omp = 1 if omp else 0
        self.h = self.L.ko_create(ctypes.byref(inp))
        if not self.h:
            raise ValueError("oracle rejected input: " + self.L.ko_error().decode())

    def __del__(self):
        if getattr(self, "h", None):
            self.L.ko_destroy(self.h)
            self.h = None

    @property
    def num_sdoms(self):
        return self.L.ko_num_sdoms(self.h)

    @property
    def num_pr(self):
        return self.L.ko_num_pr(self.h)

    def sdom(self, s):
        buf = (ctypes.c_int64 * 26)()
        if self.L.ko_sdom_info(self.h, s, buf) != 0:
            raise IndexError(s)
        return dict(zip(SDOM_KEYS, list(buf)))

    def materials(self):
        st = (ctypes.c_double * 3)()
        ss = (ctypes.c_double * 3)()
        self.L.ko_materials(self.h, st, ss)
        return np.array(st[:]), np.array(ss[:])

    def get(self, name, chunk=0):
        n = self.L.ko_get(self.h, name.encode(), chunk, None)
        if n < 0:
            raise KeyError((name, chunk))
        out = np.empty(n, dtype=np.float64)
        self.L.ko_get(self.h, name.encode(), chunk, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        return out

    def set(self, name, chunk, values):
        v = np.ascontiguousarray(values, dtype=np.float64)
        n = self.L.ko_set(self.h, name.encode(), chunk, v.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        if n < 0:
            raise KeyError((name, chunk))

    def kconst(self, name, value=0.0):
        self.L.ko_kconst_all(self.h, name.encode(), value)

    def ltimes(self):
        self.L.ko_ltimes(self.h, self.omp)

    def lplustimes(self):
        self.L.ko_lplustimes(self.h, self.omp)

    def scattering(self):
        self.L.ko_scattering(self.h, self.omp)

    def source(self):
        self.L.ko_source(self.h, self.omp)

    def sweep_subdomain(self, s):
        self.L.ko_sweep_subdomain(self.h, s, self.omp)

    def sweep_solver(self, block_jacobi=False):
        if block_jacobi:
            self.L.ko_sweep_solver_bj(self.h, self.omp)
        else:
            self.L.ko_sweep_solver(self.h, self.omp)

    def population(self):
        a, b = ctypes.c_double(), ctypes.c_double()
        self.L.ko_population(self.h, self.omp, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def init_psi(self):
        self.L.ko_init_psi(self.h)

    def iterate(self):
        a, b = ctypes.c_double(), ctypes.c_double()
        self.L.ko_iterate(self.h, self.omp, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def timer(self, name):
        c = ctypes.c_int64()
        t = self.L.ko_timer(self.h, name.encode(), ctypes.byref(c))
        return t, c.value

    def timer_reset(self):
        self.L.ko_timer_reset(self.h)


def omp_threads():
    return lib("ref").ko_omp_threads()
