"""torch.autograd.gradcheck of kripke_amd.autograd.SourceIterations (FP64,
tiny problem).  Run as its own process: torch must be imported before
libkripke_amd.so is loaded, so that both use torch's HIP runtime."""
import os
import sys

import torch  # noqa: E402  (first, see above)

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

import kripke_amd as ka  # noqa: E402
from kripke_amd.adjoint import DifferentiableSolver, chunk_get  # noqa: E402
from kripke_amd.autograd import SourceIterations  # noqa: E402


def main():
    if not torch.cuda.is_available():
        print("torch sees no device")
        return 2
    ctx = ka.Context(0)
    inp = ka.input_vars(zones=(6, 6, 6), groups=2, quad=(4, 2), legendre=1, layout="DGZ", dset=8, gset=1,
                        zset=(2, 1, 1), geometry=1, seed=20261202)
    sv = DifferentiableSolver(ctx, inp, iterations=2, checkpoint_every=1)
    sigt = [chunk_get(ctx, p, n) for p, n in zip(sv.param_ptrs("sigt"), sv.nparam)]
    # well-scaled parameters (sigma_t in [0.5, 1.5]): gradcheck's absolute eps is a small relative step
    t = torch.tensor(np.stack(sigt) + 0.5, dtype=torch.float64, device="cuda", requires_grad=True)
    s = (0.6 * t.detach()).clone().requires_grad_(True)
    ok = torch.autograd.gradcheck(lambda a, b: SourceIterations.apply(a, b, sv), (t, s), eps=1e-6, atol=1e-8,
                                  rtol=1e-6, fast_mode=True)
    # and a loss that reaches both parameters through phi: one backward by hand
    phi = SourceIterations.apply(t, s, sv)
    loss = (phi ** 2).sum()
    loss.backward()
    finite = bool(torch.isfinite(t.grad).all() and torch.isfinite(s.grad).all())
    print("gradcheck", ok, "finite grads", finite)
    sv.close()
    ctx.close()
    return 0 if (ok and finite) else 1


if __name__ == "__main__":
    sys.exit(main())
