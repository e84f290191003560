"""torch.autograd.Function over the differentiable source iteration.

    phi_K = SourceIterations.apply(sigt, s, solver)

sigt and s are CUDA float64 tensors of shape (num_pr, n) holding sigma_t and
the per-zone isotropic scattering cross section of every PR chunk in
sigt_zonal layout; phi_K is (num_pr, m) in phi's layout.  The HIP kernels run
on the library's own stream; the torch stream and the library stream are
synchronised at the boundary of each call.  torch is only the carrier of the
tensors here: the forward iterations, the adjoint chain and the gradients
are the library's kernels (kripke_amd/adjoint.py).

Import torch before the library is first loaded (before any Context is
created): torch ships its own HIP runtime, and the library then binds to that
same runtime instead of bringing up a second one.
"""
import torch

from .adjoint import DifferentiableSolver, _copy


def _check_runtime():
    if not torch.cuda.is_available():
        raise RuntimeError("torch sees no HIP device: import torch before the first kripke_amd Context is "
                           "created, so that the library and torch share one HIP runtime")


def _ptrs(t):
    return [t[c].data_ptr() for c in range(t.shape[0])]


class SourceIterations(torch.autograd.Function):
    @staticmethod
    def forward(ctx, sigt, s, solver):
        _check_runtime()
        if sigt.dtype != torch.float64 or s.dtype != torch.float64 or not sigt.is_cuda or not s.is_cuda:
            raise TypeError("sigt and s must be CUDA float64 tensors")
        sigt, s = sigt.contiguous(), s.contiguous()
        if list(sigt.shape) != [solver.npr, solver.nparam[0]] or sigt.shape != s.shape:
            raise ValueError("sigt and s must have shape (num_pr, groups*zones)")
        torch.cuda.synchronize()
        solver.set_params(_ptrs(sigt), _ptrs(s))
        solver.forward()
        out = torch.empty((solver.npr, solver.nphi[0]), dtype=torch.float64, device=sigt.device)
        for c, p in enumerate(solver.phi_ptrs()):
            _copy(solver.ctx, out[c].data_ptr(), p, solver.nphi[c])
        solver.synchronize()
        ctx.solver = solver
        ctx.save_for_backward(sigt, s)
        return out

    @staticmethod
    def backward(ctx, grad_phi):
        solver = ctx.solver
        sigt, s = ctx.saved_tensors
        grad_phi = grad_phi.contiguous()
        torch.cuda.synchronize()
        solver.set_params(_ptrs(sigt), _ptrs(s))  # the solver may have run other parameters since
        solver.forward()
        solver.backward(_ptrs(grad_phi))
        gt = torch.empty_like(sigt)
        gs = torch.empty_like(s)
        for c in range(solver.npr):
            _copy(solver.ctx, gt[c].data_ptr(), solver.param_ptrs("grad_sigt")[c], solver.nparam[c])
            _copy(solver.ctx, gs[c].data_ptr(), solver.param_ptrs("grad_s")[c], solver.nparam[c])
        solver.synchronize()
        return gt, gs, None


def source_iterations(sigt, s, solver):
    return SourceIterations.apply(sigt, s, solver)


__all__ = ["SourceIterations", "source_iterations", "DifferentiableSolver"]
