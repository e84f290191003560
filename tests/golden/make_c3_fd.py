"""Writes c3_oracle_fd.npz: central differences of the CPU oracle's forward
run (reference arithmetic, kind "ref") at the sigma_t points the GPU gradient
test checks, on the differentiable workload of tests/c3_problem.py.

The oracle runs the reference's own iteration (SteadyStateSolver.cpp:44-104:
LTimes, scattering with the generated sigma_s, source, LPlusTimes, sweep)
10 times from psi = 0 with its sigt_zonal field set to the parameters, then
one more LTimes (phi_K = L psi_K).  Only sigma_t points are differenced here:
the oracle's scattering is the reference's per-material one, which has no
per-zone parameter to perturb.

Stored: the checked entries and steps, J at the base point, J(+h), J(-h),
the differences, and the first sigma_t values as a check that the test and
this script saw the same generated field.

    python tests/golden/make_c3_fd.py        (about 10 minutes on 8 cores)
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "tests"), os.path.join(ROOT, "oracle"), ROOT]

import c3_problem as c3  # noqa: E402
import pyoracle as po  # noqa: E402


def phi_k(o, sigt):
    o.set("sigt_zonal", 0, sigt)
    o.init_psi()
    for _ in range(c3.ITERATIONS):
        o.iterate()
    o.kconst("phi", 0.0)
    o.ltimes()
    return o.get("phi", 0)


def main():
    t0 = time.time()
    o = po.Oracle(po.make_input(**c3.KW), kind="ref", omp=True)
    base = o.get("sigt_zonal", 0)
    target = phi_k(o, base * c3.target_scale())
    J0 = c3.loss(phi_k(o, base), target)
    print("J0 %.17g (%.1f s)" % (J0, time.time() - t0), flush=True)
    entries, steps, jp, jm = [], [], [], []
    for which, e in c3.parameters():
        if which != 0:
            continue
        h = c3.H_REL * abs(base[e])
        vals = []
        for sgn in (1.0, -1.0):
            v = base.copy()
            v[e] += sgn * h
            vals.append(c3.loss(phi_k(o, v), target))
        entries.append(e)
        steps.append(h)
        jp.append(vals[0])
        jm.append(vals[1])
        print("entry %d: fd %.12e (%.1f s)" % (e, (vals[0] - vals[1]) / (2 * h), time.time() - t0), flush=True)
    entries, steps, jp, jm = (np.array(a) for a in (entries, steps, jp, jm))
    np.savez(os.path.join(HERE, "c3_oracle_fd.npz"), entries=entries, steps=steps, J0=np.array(J0), Jp=jp, Jm=jm,
             fd=(jp - jm) / (2 * steps), sigt_head=base[:64], omp_threads=np.array(po.omp_threads()))
    print("wrote c3_oracle_fd.npz in %.1f s" % (time.time() - t0))


if __name__ == "__main__":
    main()
