"""kripke_amd — MI355X-native source-iteration hot path of LLNL/Kripke.

The product is libkripke_amd.so (HIP kernels for gfx950 + C++ host mirror)
behind the C-ABI in include/kripke_amd.h; this package only binds it.
"""
from .capi import (LAYOUTS, Batch, Context, HostProblem, InputVars, KernelProblem, KripkeError,  # noqa: F401
                   autotune_nesting, device_count, input_vars, lib)
