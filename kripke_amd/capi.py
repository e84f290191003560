"""ctypes binding of libkripke_amd.so (include/kripke_amd.h).

Loads the in-tree library and fails loudly when it is missing: there is no
CPU fallback behind this module.
"""
import atexit
import ctypes
import weakref
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libkripke_amd.so")

LAYOUTS = ["DGZ", "DZG", "GDZ", "GZD", "ZDG", "ZGD"]

KR_OK = 0
ERRORS = {1: "KR_ERR_INVALID", 2: "KR_ERR_HIP", 3: "KR_ERR_NOMEM", 4: "KR_ERR_UNSUPPORTED", 5: "KR_ERR_NO_DEVICE"}


class KripkeError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s: %s" % (ERRORS.get(code, code), msg))
        self.code = code


class InputVars(ctypes.Structure):
    """kr_input_vars."""

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


def input_vars(zones=(16, 16, 16), groups=32, quad=(8, 12), directions=None, legendre=4, layout="DGZ",
               dset=8, gset=1, zset=(1, 1, 1), geometry=1, seed=20261202, subsamples=4,
               sigt=(0.1, 0.0001, 0.1), sigs=(0.05, 0.00005, 0.05), source_strength=1.0, pmethod="sweep"):
    """InputVariables with the reference's meaning (src/Kripke/InputVariables.cpp:17-46).
    quad=(polar, azimuthal) is the Gauss-Legendre product set; quad=None the
    dummy S2 set with `directions` points; pmethod is --pmethod, "sweep" or
    "bj" (src/kripke.cpp:329-340)."""
    v = InputVars()
    v.nx, v.ny, v.nz = zones
    if quad is None:
        v.num_directions = directions
        v.quad_num_polar = v.quad_num_azimuthal = 0
    else:
        v.quad_num_polar, v.quad_num_azimuthal = quad
        v.num_directions = quad[0] * quad[1]
    v.num_groups = groups
    v.legendre_order = legendre
    v.layout = LAYOUTS.index(layout) if isinstance(layout, str) else int(layout)
    v.num_dirsets = dset
    v.num_groupsets = gset
    for a in range(3):
        v.num_zonesets[a] = zset[a]
    v.num_material_subsamples = subsamples
    v.geometry = geometry
    v.seed = seed
    for m in range(3):
        v.sigt[m] = sigt[m]
        v.sigs[m] = sigs[m]
    v.source_strength = source_strength
    if pmethod not in ("sweep", "bj"):
        raise ValueError("pmethod must be 'sweep' or 'bj'")
    v.parallel_method = 1 if pmethod == "bj" else 0
    return v


# exported entry points, (name, restype, argtypes); the ABI test checks this
# list against include/kripke_amd.h
P = ctypes.c_void_p
DP = ctypes.POINTER(ctypes.c_double)
SIGNATURES = [
    ("kr_context_create", ctypes.c_int, [ctypes.c_int32, ctypes.POINTER(P)]),
    ("kr_context_destroy", ctypes.c_int, [P]),
    ("kr_last_error", ctypes.c_char_p, []),
    ("kr_device_count", ctypes.c_int, [ctypes.POINTER(ctypes.c_int32)]),
    ("kr_malloc", ctypes.c_int, [P, ctypes.c_size_t, ctypes.POINTER(P)]),
    ("kr_free", ctypes.c_int, [P, P]),
    ("kr_memcpy_htod", ctypes.c_int, [P, P, P, ctypes.c_size_t]),
    ("kr_memcpy_dtoh", ctypes.c_int, [P, P, P, ctypes.c_size_t]),
    ("kr_synchronize", ctypes.c_int, [P]),
    ("kr_problem_create", ctypes.c_int, [P, P, ctypes.POINTER(P)]),
    ("kr_problem_destroy", ctypes.c_int, [P]),
    ("kr_kconst", ctypes.c_int, [P, P, ctypes.c_int64, ctypes.c_double]),
    ("kr_kcopy", ctypes.c_int, [P, P, P, ctypes.c_int64]),
    ("kr_ltimes", ctypes.c_int, [P, ctypes.c_int32]),
    ("kr_lplustimes", ctypes.c_int, [P, ctypes.c_int32]),
    ("kr_scattering", ctypes.c_int, [P, ctypes.c_int32]),
    ("kr_source", ctypes.c_int, [P, ctypes.c_double]),
    ("kr_sweep_subdomain", ctypes.c_int, [P, ctypes.c_int32]),
    ("kr_sweep_solver", ctypes.c_int, [P, ctypes.POINTER(ctypes.c_int32), ctypes.c_int32]),
    ("kr_sweep_solver_bj", ctypes.c_int, [P, ctypes.POINTER(ctypes.c_int32), ctypes.c_int32]),
    ("kr_sweep_solver_population", ctypes.c_int, [P, ctypes.POINTER(ctypes.c_int32), ctypes.c_int32, DP]),
    ("kr_population", ctypes.c_int, [P, DP]),
    ("kr_sweep_solver_population_noalloc", ctypes.c_int, [P, ctypes.POINTER(ctypes.c_int32), ctypes.c_int32, DP]),
    ("kr_ltimes_adjoint", ctypes.c_int, [P]),
    ("kr_lplustimes_adjoint", ctypes.c_int, [P]),
    ("kr_scattering_adjoint", ctypes.c_int, [P, ctypes.c_int32]),
    ("kr_scattering_zonal", ctypes.c_int, [P, ctypes.POINTER(P), ctypes.c_int32]),
    ("kr_scattering_zonal_adjoint", ctypes.c_int, [P, ctypes.POINTER(P), ctypes.c_int32]),
    ("kr_sweep_solver_adjoint", ctypes.c_int, [P]),
    ("kr_sweep_solver_adjoint_seeded", ctypes.c_int, [P]),
    ("kr_sdom_exits", ctypes.c_int, [P, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32)]),
    ("kr_grad_sigt", ctypes.c_int, [P, P, ctypes.POINTER(P), ctypes.c_int32]),
    ("kr_grad_sigs_zonal", ctypes.c_int, [P, P, ctypes.POINTER(P), ctypes.c_int32]),
    ("kr_diff_create", ctypes.c_int, [P, P, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_double,
                                      ctypes.POINTER(P)]),
    ("kr_diff_destroy", ctypes.c_int, [P]),
    ("kr_diff_num_exits", ctypes.c_int, [P, ctypes.POINTER(ctypes.c_int32)]),
    ("kr_diff_exit", ctypes.c_int, [P, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32),
                                    ctypes.POINTER(ctypes.c_int64)]),
    ("kr_diff_set_params", ctypes.c_int, [P, ctypes.c_int32, ctypes.POINTER(P), ctypes.POINTER(P)]),
    ("kr_diff_forward", ctypes.c_int, [P, ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(P), ctypes.POINTER(P)]),
    ("kr_diff_backward", ctypes.c_int, [P, ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(P), ctypes.POINTER(P),
                                        ctypes.POINTER(P), ctypes.POINTER(P)]),
    ("kr_problem_autotune", ctypes.c_int, [P, ctypes.POINTER(ctypes.c_int32)]),
    ("kr_problem_set_tuning", ctypes.c_int, [P, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]),
    ("kr_problem_sweep_kind", ctypes.c_int, [P, ctypes.POINTER(ctypes.c_int32)]),
    ("kr_batch_create", ctypes.c_int, [P, ctypes.POINTER(P), DP, ctypes.c_int32, ctypes.POINTER(P)]),
    ("kr_batch_destroy", ctypes.c_int, [P]),
    ("kr_batch_iterate", ctypes.c_int, [P, ctypes.c_int32, DP]),
    ("kr_problem_set_timing", ctypes.c_int, [P, ctypes.c_int32]),
    ("kr_problem_timer", ctypes.c_int, [P, ctypes.c_char_p, DP, ctypes.POINTER(ctypes.c_int64)]),
    ("kr_problem_timer_reset", ctypes.c_int, [P]),
    ("kr_host_create", ctypes.c_int, [P, ctypes.POINTER(InputVars), ctypes.POINTER(P)]),
    ("kr_host_destroy", ctypes.c_int, [P]),
    ("kr_host_num_sdoms", ctypes.c_int, [P, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32),
                                         ctypes.POINTER(ctypes.c_int32)]),
    ("kr_host_kernel_problem", P, [P]),
    ("kr_host_field_get", ctypes.c_int, [P, ctypes.c_char_p, ctypes.c_int32, DP, ctypes.POINTER(ctypes.c_int64)]),
    ("kr_host_field_set", ctypes.c_int, [P, ctypes.c_char_p, ctypes.c_int32, DP]),
    ("kr_host_field_device", ctypes.c_int, [P, ctypes.c_char_p, ctypes.c_int32, ctypes.POINTER(P),
                                             ctypes.POINTER(ctypes.c_int64)]),
    ("kr_host_init_psi", ctypes.c_int, [P]),
    ("kr_host_iterate", ctypes.c_int, [P, ctypes.c_int32, DP]),
    ("kr_host_solve", ctypes.c_int, [P, ctypes.c_int32, ctypes.c_int32, DP]),
    ("kr_host_autotune_nesting", ctypes.c_int, [P, ctypes.POINTER(InputVars), ctypes.POINTER(ctypes.c_int32), DP]),
]

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libkripke_amd.so is not built (run __graft_entry__.build() or "
                              "python kripke_amd/build.py); there is no fallback path")
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc, what=""):
    if rc != KR_OK:
        raise KripkeError(rc, "%s: %s" % (what, lib().kr_last_error().decode()))


def device_count():
    n = ctypes.c_int32()
    check(lib().kr_device_count(ctypes.byref(n)), "kr_device_count")
    return n.value


def autotune_nesting(ctx, inp):
    """kr_host_autotune_nesting: the fastest of the six nestings for the
    problem `inp` on ctx (each built, sweep-tuned and timed for one source
    iteration); returns (layout name, ms per iteration).  Cached per context
    and shape."""
    nest, ms = ctypes.c_int32(), ctypes.c_double()
    check(lib().kr_host_autotune_nesting(ctx.h, ctypes.byref(inp), ctypes.byref(nest), ctypes.byref(ms)),
          "kr_host_autotune_nesting")
    return LAYOUTS[nest.value], ms.value


_live_contexts = weakref.WeakSet()


@atexit.register
def _close_live_contexts():
    """Contexts still open at interpreter exit are closed (problems first)
    while the HIP runtime is still up, not from objects' finalizers in
    whatever order the interpreter tears modules down."""
    for ctx in list(_live_contexts):
        try:
            ctx.close()
        except Exception:
            pass


class Context:
    """kr_context: one device, one HIP stream."""

    def __init__(self, device=0):
        self.h = P()
        self._children = weakref.WeakSet()  # problems built on this context
        check(lib().kr_context_create(device, ctypes.byref(self.h)), "kr_context_create")
        _live_contexts.add(self)

    def synchronize(self):
        check(lib().kr_synchronize(self.h), "kr_synchronize")

    def close(self):
        """Destroys the problems still alive on this context, then the context."""
        if self.h:
            for child in list(self._children):
                child.close()
            lib().kr_context_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class KernelProblem:
    """The registered kr_problem of a host problem: the Kernel:: entry points."""

    def __init__(self, handle, owner):
        self._h = handle
        self.owner = owner  # keeps the host problem alive

    @property
    def h(self):
        if not self.owner.h:
            raise KripkeError(1, "the host problem of this kernel problem is closed")
        return self._h

    def ltimes(self, accumulate=True):
        check(lib().kr_ltimes(self.h, 1 if accumulate else 0), "kr_ltimes")

    def lplustimes(self, accumulate=True):
        check(lib().kr_lplustimes(self.h, 1 if accumulate else 0), "kr_lplustimes")

    def scattering(self, accumulate=True):
        check(lib().kr_scattering(self.h, 1 if accumulate else 0), "kr_scattering")

    def source(self, strength=1.0):
        check(lib().kr_source(self.h, strength), "kr_source")

    def sweep_subdomain(self, sdom):
        check(lib().kr_sweep_subdomain(self.h, sdom), "kr_sweep_subdomain")

    def sweep_solver(self, sdoms, block_jacobi=False):
        arr = (ctypes.c_int32 * len(sdoms))(*sdoms)
        if block_jacobi:
            check(lib().kr_sweep_solver_bj(self.h, arr, len(sdoms)), "kr_sweep_solver_bj")
        else:
            check(lib().kr_sweep_solver(self.h, arr, len(sdoms)), "kr_sweep_solver")

    def sweep_solver_population(self, sdoms):
        arr = (ctypes.c_int32 * len(sdoms))(*sdoms)
        v = ctypes.c_double()
        check(lib().kr_sweep_solver_population(self.h, arr, len(sdoms), ctypes.byref(v)),
              "kr_sweep_solver_population")
        return v.value

    def sweep_solver_population_noalloc(self, sdoms):
        """kr_sweep_solver_population_noalloc: the same with the partials in
        the device's static scratch array (no device allocation)."""
        arr = (ctypes.c_int32 * len(sdoms))(*sdoms)
        v = ctypes.c_double()
        check(lib().kr_sweep_solver_population_noalloc(self.h, arr, len(sdoms), ctypes.byref(v)),
              "kr_sweep_solver_population_noalloc")
        return v.value

    def population(self):
        v = ctypes.c_double()
        check(lib().kr_population(self.h, ctypes.byref(v)), "kr_population")
        return v.value

    # ---- transposed operators and parameter gradients (include/kripke_amd.h)
    def ltimes_adjoint(self):
        check(lib().kr_ltimes_adjoint(self.h), "kr_ltimes_adjoint")

    def lplustimes_adjoint(self):
        check(lib().kr_lplustimes_adjoint(self.h), "kr_lplustimes_adjoint")

    def scattering_adjoint(self, accumulate=False):
        check(lib().kr_scattering_adjoint(self.h, 1 if accumulate else 0), "kr_scattering_adjoint")

    def scattering_zonal(self, ptrs, accumulate=False, adjoint=False):
        arr = (P * len(ptrs))(*ptrs)
        fn = lib().kr_scattering_zonal_adjoint if adjoint else lib().kr_scattering_zonal
        check(fn(self.h, arr, 1 if accumulate else 0), "kr_scattering_zonal")

    def sweep_solver_adjoint(self, seeded=False):
        """rhs = H^-T psi; seeded: with adjoint inflow on the domain boundary
        read from the own plane chunks (kr_sweep_solver_adjoint_seeded)."""
        if seeded:
            check(lib().kr_sweep_solver_adjoint_seeded(self.h), "kr_sweep_solver_adjoint_seeded")
        else:
            check(lib().kr_sweep_solver_adjoint(self.h), "kr_sweep_solver_adjoint")

    def sdom_exits(self, sdom):
        """(i, j, k) flags: the subdomain's downwind face on that axis is the
        domain boundary."""
        e = (ctypes.c_int32 * 3)()
        check(lib().kr_sdom_exits(self.h, int(sdom), e), "kr_sdom_exits")
        return tuple(bool(v) for v in e)

    def grad_sigt(self, adj, ptrs, accumulate=False):
        arr = (P * len(ptrs))(*ptrs)
        check(lib().kr_grad_sigt(self.h, adj.h, arr, 1 if accumulate else 0), "kr_grad_sigt")

    def grad_sigs_zonal(self, adj, ptrs, accumulate=False):
        arr = (P * len(ptrs))(*ptrs)
        check(lib().kr_grad_sigs_zonal(self.h, adj.h, arr, 1 if accumulate else 0), "kr_grad_sigs_zonal")

    def autotune(self):
        """Pick the fastest sweep shape for this problem (overwrites psi and the
        planes); returns (kernel kind, zones per lane or patch columns, unit %)."""
        c = (ctypes.c_int32 * 3)()
        check(lib().kr_problem_autotune(self.h, c), "kr_problem_autotune")
        return tuple(c)

    def set_tuning(self, v_max=0, cap_div=1, r_scale_pct=100):
        """Install one autotuner candidate (kr_problem_set_tuning); returns the
        sweep kernel kind now installed (0 systolic, 1 row scan, 3 wave, 4
        systolic with the LDS k-face ring)."""
        check(lib().kr_problem_set_tuning(self.h, v_max, cap_div, r_scale_pct), "kr_problem_set_tuning")
        k = ctypes.c_int32()
        check(lib().kr_problem_sweep_kind(self.h, ctypes.byref(k)), "kr_problem_sweep_kind")
        return k.value

    def set_timing(self, on=True):
        check(lib().kr_problem_set_timing(self.h, 1 if on else 0), "kr_problem_set_timing")

    def timer(self, name):
        s, c = ctypes.c_double(), ctypes.c_int64()
        check(lib().kr_problem_timer(self.h, name.encode(), ctypes.byref(s), ctypes.byref(c)), "kr_problem_timer")
        return s.value, c.value

    def timer_reset(self):
        check(lib().kr_problem_timer_reset(self.h), "kr_problem_timer_reset")


class Batch:
    """kr_batch: independent host problems on one context advanced by one
    hipGraph launch per call (include/kripke_amd.h)."""

    def __init__(self, ctx, host_problems):
        self.ctx, self.problems = ctx, list(host_problems)
        ks = [hp.kernels() for hp in self.problems]
        arr = (P * len(ks))(*[k.h for k in ks])
        st = (ctypes.c_double * len(ks))(*[hp.inp.source_strength for hp in self.problems])
        self.h = P()
        check(lib().kr_batch_create(ctx.h, arr, st, len(ks), ctypes.byref(self.h)), "kr_batch_create")
        ctx._children.add(self)  # destroyed with the context

    def iterate(self, niter):
        """Populations, shape (problems, niter)."""
        out = np.zeros(len(self.problems) * niter, dtype=np.float64)
        check(lib().kr_batch_iterate(self.h, niter, out.ctypes.data_as(DP)), "kr_batch_iterate")
        return out.reshape(len(self.problems), niter)

    def close(self):
        if self.h:
            lib().kr_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class HostProblem:
    """generateProblem + DataStore of the host mirror (kr_host_*).
    ctx=None builds host-side setup fields only."""

    def __init__(self, inp, ctx=None):
        self.ctx = ctx
        if ctx is not None:
            if not ctx.h:
                raise KripkeError(1, "context is closed")
            ctx._children.add(self)
        self.inp = inp
        self.h = P()
        check(lib().kr_host_create(ctx.h if ctx is not None else None, ctypes.byref(inp), ctypes.byref(self.h)),
              "kr_host_create")
        ns, npr, nr = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        check(lib().kr_host_num_sdoms(self.h, ctypes.byref(ns), ctypes.byref(npr), ctypes.byref(nr)), "num_sdoms")
        self.num_sdoms, self.num_pr, self.num_r = ns.value, npr.value, nr.value

    def kernels(self):
        kh = lib().kr_host_kernel_problem(self.h)
        if not kh:
            raise KripkeError(1, "host problem has no device problem")
        return KernelProblem(kh, self)

    def get(self, name, chunk=0):
        n = ctypes.c_int64()
        check(lib().kr_host_field_get(self.h, name.encode(), chunk, None, ctypes.byref(n)), "field_get " + name)
        out = np.empty(n.value, dtype=np.float64)
        check(lib().kr_host_field_get(self.h, name.encode(), chunk, out.ctypes.data_as(DP), ctypes.byref(n)),
              "field_get " + name)
        return out

    def set(self, name, chunk, values):
        v = np.ascontiguousarray(values, dtype=np.float64)
        check(lib().kr_host_field_set(self.h, name.encode(), chunk, v.ctypes.data_as(DP)), "field_set " + name)

    def device_ptr(self, name, chunk=0):
        """(device address, element count) of a double field's chunk."""
        ptr, n = P(), ctypes.c_int64()
        check(lib().kr_host_field_device(self.h, name.encode(), chunk, ctypes.byref(ptr), ctypes.byref(n)),
              "field_device " + name)
        return ptr.value, n.value

    def init_psi(self):
        check(lib().kr_host_init_psi(self.h), "init_psi")

    def iterate(self, mode=1):
        v = ctypes.c_double()
        check(lib().kr_host_iterate(self.h, mode, ctypes.byref(v)), "iterate")
        return v.value

    def solve(self, niter, mode=1):
        out = np.zeros(niter, dtype=np.float64)
        check(lib().kr_host_solve(self.h, niter, mode, out.ctypes.data_as(DP)), "solve")
        return out

    def close(self):
        if self.h:
            lib().kr_host_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
