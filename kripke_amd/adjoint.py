"""Reverse-mode gradient of K source iterations (no reference counterpart).

The forward model is K steady-state iterations from psi = 0
(src/Kripke/SteadyStateSolver.cpp:44-104) with the per-zone parameters
made explicit:

    phi_n      = L psi_n
    phi_out_n  = S(s) phi_n + q            per-zone isotropic scattering s(g,z)
    rhs_n      = L+ phi_out_n
    psi_{n+1}  = H(sigt)^-1 rhs_n          the diamond-difference sweep
    phi_K      = L psi_K                   the output

with sigt(g,z) the sigma_t field (sigt_zonal) and s(g,z) the group-diagonal
n = 0 scattering cross section, both one array per PR chunk in sigt_zonal
layout.  The forward pass, the phi checkpoints, the recomputation between
them and the adjoint chain all run inside the library's C-ABI pair
kr_diff_forward / kr_diff_backward (include/kripke_amd.h); this module owns
the two problems the pair runs on and passes device addresses.
"""
import ctypes

import numpy as np

from .capi import HostProblem, check, lib

P = ctypes.c_void_p


class DeviceBuffer:
    """One kr_malloc'd device array of doubles."""

    def __init__(self, ctx, n):
        self.ctx, self.n = ctx, int(n)
        self.ptr = ctypes.c_void_p()
        check(lib().kr_malloc(ctx.h, max(self.n, 1) * 8, ctypes.byref(self.ptr)), "kr_malloc")

    def free(self):
        if self.ptr:
            lib().kr_free(self.ctx.h, self.ptr)
            self.ptr = ctypes.c_void_p()


def _copy(ctx, dst, src, n):
    check(lib().kr_kcopy(ctx.h, ctypes.c_void_p(dst), ctypes.c_void_p(src), int(n)), "kr_kcopy")


def _fill(ctx, dst, n, v):
    check(lib().kr_kconst(ctx.h, ctypes.c_void_p(dst), int(n), float(v)), "kr_kconst")


def _ptr_array(ptrs):
    """Host array of device addresses (None entries stay NULL); None -> NULL."""
    if ptrs is None:
        return None
    return (P * len(ptrs))(*[p if p else None for p in ptrs])


PLANES = ("i_plane", "j_plane", "k_plane")


class DifferentiableSolver:
    """K source iterations as a function of (sigt, s) with phi checkpoints
    every `checkpoint_every` iterations, for `fields` independent parameter
    sets (kr_diff)."""

    def __init__(self, ctx, inp, iterations, checkpoint_every=2, fields=1):
        if iterations < 1 or checkpoint_every < 1 or fields < 1:
            raise ValueError("iterations, checkpoint_every and fields must be >= 1")
        self.ctx, self.K, self.ck, self.F = ctx, int(iterations), int(checkpoint_every), int(fields)
        self.fwd = HostProblem(inp, ctx)
        self.adj = HostProblem(inp, ctx)
        self.kf, self.ka = self.fwd.kernels(), self.adj.kernels()
        self.npr, self.nsdom = self.fwd.num_pr, self.fwd.num_sdoms
        # per PR chunk: element counts of the sigma_t and phi chunks
        self.nparam = [self.fwd.device_ptr("sigt_zonal", c)[1] for c in range(self.npr)]
        self.nphi = [self.fwd.device_ptr("phi", c)[1] for c in range(self.npr)]
        self.h = P()
        check(lib().kr_diff_create(self.kf.h, self.ka.h, self.K, self.ck, self.F, float(inp.source_strength),
                                   ctypes.byref(self.h)), "kr_diff_create")
        ne = ctypes.c_int32()
        check(lib().kr_diff_num_exits(self.h, ctypes.byref(ne)), "kr_diff_num_exits")
        self.exits = []  # (sdom, plane name, element count), kr_diff's exit order
        for e in range(ne.value):
            sd, ax, n = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int64()
            check(lib().kr_diff_exit(self.h, e, ctypes.byref(sd), ctypes.byref(ax), ctypes.byref(n)), "kr_diff_exit")
            self.exits.append((sd.value, PLANES[ax.value], n.value))
        self._gt = [DeviceBuffer(ctx, n) for n in self.nparam]  # gradients of backward()
        self._gs = [DeviceBuffer(ctx, n) for n in self.nparam]
        ctx._children.add(self)  # released before its context

    # ---- parameters ------------------------------------------------------
    def set_params(self, sigt_ptrs, s_ptrs, field=0):
        """Device arrays (one per PR chunk, sigt_zonal layout) of sigma_t and
        s for one field; copied before the call returns."""
        check(lib().kr_diff_set_params(self.h, int(field), _ptr_array(sigt_ptrs), _ptr_array(s_ptrs)),
              "kr_diff_set_params")

    def param_ptrs(self, which):
        """Device addresses of the forward problem's sigma_t ('sigt': the
        generated field until a forward pass writes a field's own) or of the
        gradients of the last backward() ('grad_sigt', 'grad_s')."""
        if which == "sigt":
            return [self.fwd.device_ptr("sigt_zonal", c)[0] for c in range(self.npr)]
        bufs = {"grad_sigt": self._gt, "grad_s": self._gs}[which]
        return [b.ptr.value for b in bufs]

    # ---- forward ---------------------------------------------------------
    def forward(self, field=0):
        """Runs the K iterations of one field from psi = 0; phi_K is left in the
        forward problem's phi (phi_ptrs()) and the last sweep's exiting faces
        in its plane chunks (exit_faces())."""
        self.forward_fields(field, 1)

    def forward_fields(self, first, num, phi_k=None, faces=None):
        """Fields [first, first+num) in one call.  phi_k: num*num_pr device
        addresses (field-major) receiving each field's phi_K; faces:
        num*len(exits) addresses receiving its exiting faces (either None)."""
        check(lib().kr_diff_forward(self.h, int(first), int(num), _ptr_array(phi_k), _ptr_array(faces)),
              "kr_diff_forward")

    def phi_ptrs(self, problem="fwd"):
        hp = self.fwd if problem == "fwd" else self.adj
        return [hp.device_ptr("phi", c)[0] for c in range(self.npr)]

    def exit_faces(self):
        """(sdom, plane name, device address, count) of every plane chunk of
        the forward problem that holds faces leaving the domain; after
        forward() they hold the last sweep's exiting angular flux."""
        return [(s, name, self.fwd.device_ptr(name, s)[0], n) for s, name, n in self.exits]

    # ---- backward --------------------------------------------------------
    def backward(self, phibar_ptrs=None, facebar=None, field=0):
        """Gradient of <phibar, phi_K> + <facebar, exiting faces of the last
        sweep> w.r.t. sigt and s of one field, left in param_ptrs('grad_sigt')
        and param_ptrs('grad_s').  phibar_ptrs: one device array per PR chunk
        (None: zero); facebar: {(sdom, plane name): device array} over
        exit_faces() (None: no face term)."""
        faces = None
        if facebar is not None:
            keys = [(s, name) for s, name, _ in self.exits]
            for key in facebar:
                if key not in keys:
                    raise ValueError("facebar names a plane chunk that does not leave the domain: %r" % (key,))
            faces = [facebar.get(key) for key in keys]
        self.backward_fields(field, 1, phibar_ptrs, faces, self.param_ptrs("grad_sigt"), self.param_ptrs("grad_s"))

    def backward_fields(self, first, num, phibar=None, faces=None, grad_sigt=None, grad_s=None):
        """Fields [first, first+num) in one call (flat, field-major address
        lists as forward_fields)."""
        check(lib().kr_diff_backward(self.h, int(first), int(num), _ptr_array(phibar), _ptr_array(faces),
                                     _ptr_array(grad_sigt), _ptr_array(grad_s)), "kr_diff_backward")

    def synchronize(self):
        self.ctx.synchronize()

    def close(self):
        if self.h:
            lib().kr_diff_destroy(self.h)
            self.h = None
        if self.ctx.h:
            for b in self._gt + self._gs:
                b.free()
        self.fwd.close()
        self.adj.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def chunk_get(ctx, ptr, n):
    """Device array -> numpy (synchronizes)."""
    out = np.empty(int(n), dtype=np.float64)
    check(lib().kr_memcpy_dtoh(ctx.h, out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(ptr), int(n) * 8),
          "kr_memcpy_dtoh")
    return out


def chunk_set(ctx, ptr, values):
    v = np.ascontiguousarray(values, dtype=np.float64)
    check(lib().kr_memcpy_htod(ctx.h, ctypes.c_void_p(ptr), v.ctypes.data_as(ctypes.c_void_p), v.size * 8),
          "kr_memcpy_htod")
