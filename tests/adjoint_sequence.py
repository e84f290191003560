"""The differentiable solve sequenced from Python, one C-ABI operator call
at a time (test infrastructure, not product code).

The library's kr_diff_forward/kr_diff_backward pair runs the same chain of
entry points in C++; tests hold the pair to this sequence bit for bit.  The
chain (no reference counterpart) differentiates K iterations of
src/Kripke/SteadyStateSolver.cpp:44-104 from psi = 0 with respect to per-zone
sigma_t and isotropic scattering s.
"""
from kripke_amd.adjoint import PLANES, DeviceBuffer, _copy, _fill
from kripke_amd.capi import HostProblem


class SequencedSolver:
    """K source iterations as a function of (sigt, s) with phi checkpoints
    every `checkpoint_every` iterations, sequenced call by call from Python
    over the library's operator entry points (the order kr_diff_forward and
    kr_diff_backward follow in C++)."""

    def __init__(self, ctx, inp, iterations, checkpoint_every=2):
        if iterations < 1 or checkpoint_every < 1:
            raise ValueError("iterations and checkpoint_every must be >= 1")
        self.ctx, self.K, self.ck = ctx, int(iterations), int(checkpoint_every)
        self.fwd = HostProblem(inp, ctx)
        self.adj = HostProblem(inp, ctx)
        self.kf, self.ka = self.fwd.kernels(), self.adj.kernels()
        self.npr, self.nsdom = self.fwd.num_pr, self.fwd.num_sdoms
        self.sdoms = list(range(self.nsdom))
        # per PR chunk: element counts of the sigma_t and phi chunks
        self.nparam = [self.fwd.device_ptr("sigt_zonal", c)[1] for c in range(self.npr)]
        self.nphi = [self.fwd.device_ptr("phi", c)[1] for c in range(self.npr)]
        self._s = [DeviceBuffer(ctx, n) for n in self.nparam]        # s(g,z) actually used
        self._gt = [DeviceBuffer(ctx, n) for n in self.nparam]       # gradient w.r.t. sigt
        self._gs = [DeviceBuffer(ctx, n) for n in self.nparam]       # gradient w.r.t. s
        nck = (self.K + self.ck - 1) // self.ck
        self._ckpt = [[DeviceBuffer(ctx, n) for n in self.nphi] for _ in range(nck)]
        # phi of the other iterations of one checkpoint interval, recomputed once per interval in backward()
        self._blk = [[DeviceBuffer(ctx, n) for n in self.nphi] for _ in range(self.ck - 1)]
        self.forward_done = False

    # ---- parameters ------------------------------------------------------
    def set_params(self, sigt_ptrs, s_ptrs):
        """Device arrays (one per PR chunk, sigt_zonal layout) of sigma_t and s:
        copied into both problems' sigt_zonal and into the solver's s."""
        for c in range(self.npr):
            for hp in (self.fwd, self.adj):
                _copy(self.ctx, hp.device_ptr("sigt_zonal", c)[0], sigt_ptrs[c], self.nparam[c])
            _copy(self.ctx, self._s[c].ptr.value, s_ptrs[c], self.nparam[c])
        self.forward_done = False

    def param_ptrs(self, which):
        """Device addresses of the problem-held parameters ('sigt', 's') or the
        gradients ('grad_sigt', 'grad_s')."""
        if which == "sigt":
            return [self.fwd.device_ptr("sigt_zonal", c)[0] for c in range(self.npr)]
        bufs = {"s": self._s, "grad_sigt": self._gt, "grad_s": self._gs}[which]
        return [b.ptr.value for b in bufs]

    # ---- forward ---------------------------------------------------------
    def _tail(self):
        """From fwd.phi = phi_n: phi_out, rhs and psi_{n+1} (phi left as phi_n)."""
        self.kf.scattering_zonal(self.param_ptrs("s"))
        self.kf.source(self.fwd.inp.source_strength)
        self.kf.lplustimes(accumulate=False)
        self.kf.sweep_solver(self.sdoms)

    def _step(self):
        """phi_n -> phi_{n+1} in fwd.phi."""
        self._tail()
        self.kf.ltimes(accumulate=False)

    def _set_phi(self, bufs):
        for c in range(self.npr):
            _copy(self.ctx, self.fwd.device_ptr("phi", c)[0], bufs[c].ptr.value, self.nphi[c])

    def forward(self):
        """Runs the K iterations from psi = 0; phi_K is left in the forward
        problem's phi (phi_ptrs())."""
        for c in range(self.npr):
            _fill(self.ctx, self.fwd.device_ptr("phi", c)[0], self.nphi[c], 0.0)  # phi_0 = L 0
        for n in range(self.K):
            if n % self.ck == 0:
                for c in range(self.npr):
                    _copy(self.ctx, self._ckpt[n // self.ck][c].ptr.value, self.fwd.device_ptr("phi", c)[0],
                          self.nphi[c])
            self._step()
        self.forward_done = True

    def phi_ptrs(self, problem="fwd"):
        hp = self.fwd if problem == "fwd" else self.adj
        return [hp.device_ptr("phi", c)[0] for c in range(self.npr)]

    def exit_faces(self):
        """(sdom, plane name, device address, count) of every plane chunk of
        the forward problem that holds faces leaving the domain; after
        forward() they hold the last sweep's exiting angular flux."""
        out = []
        for s in self.sdoms:
            for a, leaves in enumerate(self.kf.sdom_exits(s)):
                if leaves:
                    ptr, n = self.fwd.device_ptr(PLANES[a], s)
                    out.append((s, PLANES[a], ptr, n))
        return out

    # ---- backward --------------------------------------------------------
    def backward(self, phibar_ptrs=None, facebar=None):
        """Gradient of <phibar, phi_K> + <facebar, exiting faces of the last
        sweep> w.r.t. sigt and s, left in param_ptrs('grad_sigt') and
        param_ptrs('grad_s').  phibar_ptrs: one device array per PR chunk (None:
        zero); facebar: {(sdom, plane name): device array} over exit_faces()."""
        if not self.forward_done:
            raise RuntimeError("backward needs the forward pass of the same parameters")
        exits = {(s, name): n for s, name, _, n in self.exit_faces()}
        for key in (facebar or {}):
            if key not in exits:
                raise ValueError("facebar names a plane chunk that does not leave the domain: %r" % (key,))
        for c in range(self.npr):
            if phibar_ptrs is None:
                _fill(self.ctx, self.adj.device_ptr("phi", c)[0], self.nphi[c], 0.0)
            else:
                _copy(self.ctx, self.adj.device_ptr("phi", c)[0], phibar_ptrs[c], self.nphi[c])
            _fill(self.ctx, self._gt[c].ptr.value, self.nparam[c], 0.0)
            _fill(self.ctx, self._gs[c].ptr.value, self.nparam[c], 0.0)
        self.ka.ltimes_adjoint()                                  # psibar_K
        gt, gs = self.param_ptrs("grad_sigt"), self.param_ptrs("grad_s")
        for n in range(self.K - 1, -1, -1):
            b, c0 = n // self.ck, (n // self.ck) * self.ck
            if n == min(self.K, c0 + self.ck) - 1:
                # last iteration of an interval: recompute phi_{c0+1} .. phi_n once from the checkpoint
                self._set_phi(self._ckpt[b])
                for j in range(c0, n):
                    self._step()
                    for c in range(self.npr):
                        _copy(self.ctx, self._blk[j - c0][c].ptr.value, self.fwd.device_ptr("phi", c)[0],
                              self.nphi[c])
            else:
                self._set_phi(self._ckpt[b] if n == c0 else self._blk[n - c0 - 1])  # phi_n
            self._tail()                                          # psi_{n+1}; phi still phi_n
            if n == self.K - 1 and facebar:
                for (s, name), n_el in exits.items():             # adjoint inflow: the faces' cotangent
                    dst = self.adj.device_ptr(name, s)[0]
                    if (s, name) in facebar:
                        _copy(self.ctx, dst, facebar[(s, name)], n_el)
                    else:
                        _fill(self.ctx, dst, n_el, 0.0)
                self.ka.sweep_solver_adjoint(seeded=True)         # rhsbar_{K-1}
            else:
                self.ka.sweep_solver_adjoint()                    # rhsbar_n
            self.kf.grad_sigt(self.ka, gt, accumulate=True)
            self.ka.lplustimes_adjoint()                          # phi_outbar_n
            self.kf.grad_sigs_zonal(self.ka, gs, accumulate=True)
            self.ka.scattering_zonal(self.param_ptrs("s"), adjoint=True)  # phibar_n
            self.ka.ltimes_adjoint()                              # psibar_n

    def synchronize(self):
        self.ctx.synchronize()

    def close(self):
        for b in self._s + self._gt + self._gs:
            b.free()
        for ck in self._ckpt + self._blk:
            for b in ck:
                b.free()
        self.fwd.close()
        self.adj.close()
