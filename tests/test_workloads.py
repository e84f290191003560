"""The batched and differentiable workloads at their full sizes (GPU, through
the C-ABI).

- Batched: the 48 problems of bench.batch_configs() (zones (8..40)^3 with
  per-axis jitter, G in {8,16,32,64}, D in {32,...,128}, P1-P4, nesting and
  sweep shape autotuned), 3 iterations in one kr_batch launch: every
  problem's populations equal its own per-problem run bit for bit, and the 8
  smallest and the 2 largest problems agree with the reference-arithmetic
  CPU oracle to 1e-12 (SteadyStateSolver.cpp:44-104, Population.cpp:49-63).
- Differentiable, phi loss (tests/c3_problem.py): 32^3 zones x 32 groups x 96
  directions, P2, 10 iterations, checkpoints every 2, 64 parameters: the
  adjoint gradient against central differences of the library's forward
  run, and for the sigma_t parameters against central differences of the
  CPU oracle's forward run at the same points (tests/golden/c3_oracle_fd.npz,
  made by tests/golden/make_c3_fd.py).
- Differentiable, exiting-flux loss: 48^3 zones x 64 groups x 128 directions,
  P3, 20 iterations, checkpoints every 4, two fields in one kr_diff call (a
  truth field giving the target faces and a candidate), 8 parameters of the
  candidate against central differences of the forward run.

A central difference carries J's rounding divided by 2h; each check is
|fd - adjoint| <= 1e-6 |fd| + that noise, and the parameters are placed
where the gradient stands far above the noise.
"""
import os
import sys
import time

import numpy as np
import pytest

import kripke_amd as ka
from kripke_amd.adjoint import DeviceBuffer, DifferentiableSolver, chunk_get, chunk_set
from helpers import oracle_for

import c3_problem as c3

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def _say(msg):
    sys.stdout.write("[workloads %s] %s\n" % (time.strftime("%H:%M:%S"), msg))
    sys.stdout.flush()


# --------------------------------------------------------------------------
# batched
# --------------------------------------------------------------------------
@pytest.mark.timeout(600)
def test_batch_workload_against_per_problem_and_oracle(ctx):
    sys.path.insert(0, os.path.dirname(HERE))
    import bench
    iters = 3
    cfgs = bench.batch_configs()
    for c in cfgs:
        c["layout"] = ka.autotune_nesting(ctx, ka.input_vars(**c))[0]
    hps = [ka.HostProblem(ka.input_vars(**c), ctx) for c in cfgs]
    try:
        for hp in hps:
            hp.kernels().autotune()
            hp.init_psi()
        seq = np.array([[hp.iterate(1) for _ in range(iters)] for hp in hps])
        b = ka.Batch(ctx, hps)
        for hp in hps:
            hp.init_psi()
        pops = b.iterate(iters)
        b.close()
        _say("batch of %d problems, nestings %s" % (len(cfgs), sorted({c["layout"] for c in cfgs})))
        assert np.array_equal(pops, seq)
    finally:
        for hp in hps:
            hp.close()
    size = [c["zones"][0] * c["zones"][1] * c["zones"][2] * c["groups"] * c["quad"][0] * c["quad"][1] for c in cfgs]
    order = np.argsort(size)
    for i in list(order[:8]) + list(order[-2:]):
        kw = dict(cfgs[i], layout="DGZ")
        orf = oracle_for("ref", **kw)
        orf.init_psi()
        comp = np.array([orf.iterate()[1] for _ in range(iters)])
        rel = np.max(np.abs(pops[i] - comp) / np.abs(comp))
        _say("problem %d (%s, %s zones, G %d, D %d, P%d): population rel %.2e against the oracle" % (
            i, cfgs[i]["layout"], cfgs[i]["zones"], cfgs[i]["groups"], cfgs[i]["quad"][0] * cfgs[i]["quad"][1],
            cfgs[i]["legendre"], rel))
        assert rel <= 1e-12
        del orf


# --------------------------------------------------------------------------
# differentiable, phi loss over a detector box
# --------------------------------------------------------------------------
def _set(sv, ctx, bufs, sigt, s, field=0):
    chunk_set(ctx, bufs[0].ptr.value, sigt)
    chunk_set(ctx, bufs[1].ptr.value, s)
    sv.set_params([bufs[0].ptr.value], [bufs[1].ptr.value], field=field)


@pytest.mark.timeout(600)
def test_c3_gradient_against_finite_differences(ctx):
    fx = np.load(os.path.join(HERE, "golden", "c3_oracle_fd.npz"))
    sv = DifferentiableSolver(ctx, ka.input_vars(**c3.KW), c3.ITERATIONS, c3.CHECKPOINT)
    try:
        assert sv.npr == 1
        base = chunk_get(ctx, sv.param_ptrs("sigt")[0], sv.nparam[0])
        assert np.array_equal(base[:64], fx["sigt_head"]), "generated sigma_t differs from the oracle's"
        hp = sv.fwd
        sigs = hp.get("data/sigs").reshape(3, -1, c3.KW["groups"], c3.KW["groups"])
        s = c3.zonal_scattering(hp.get("zone_to_mixelem", 0).astype(int), hp.get("zone_to_num_mixelem", 0).astype(int),
                                hp.get("mixelem_to_material", 0).astype(int), hp.get("mixelem_to_fraction", 0),
                                np.array([np.diag(sigs[m, 0]) for m in range(3)]))
        bufs = [DeviceBuffer(ctx, base.size), DeviceBuffer(ctx, base.size)]
        wb = DeviceBuffer(ctx, sv.nphi[0])

        def phi_for(sigt, sval):
            _set(sv, ctx, bufs, sigt, sval)
            sv.forward()
            return chunk_get(ctx, sv.phi_ptrs()[0], sv.nphi[0])

        target = phi_for(base * c3.target_scale(), s)
        phi = phi_for(base, s)
        J0 = c3.loss(phi, target)
        jrel = abs(J0 - float(fx["J0"])) / float(fx["J0"])
        _say("J0 %.17g (oracle %.17g, rel %.2e)" % (J0, float(fx["J0"]), jrel))
        assert jrel <= 1e-9
        chunk_set(ctx, wb.ptr.value, np.where(c3.phi_mask(), 2.0 * (phi - target), 0.0))
        sv.backward([wb.ptr.value])
        gt = chunk_get(ctx, sv.param_ptrs("grad_sigt")[0], sv.nparam[0])
        gs = chunk_get(ctx, sv.param_ptrs("grad_s")[0], sv.nparam[0])
        oracle_fd = dict(zip(fx["entries"].tolist(), fx["fd"].tolist()))
        ratio, ratio_oracle, n_oracle = 0.0, 0.0, 0
        rel = []  # (|grad| / noise, relative error against the library's differences)
        for which, e in c3.parameters():
            vals = (base, s)[which]
            grad = (gt, gs)[which][e]
            h = c3.H_REL * abs(vals[e])
            js = []
            for sgn in (1.0, -1.0):
                v = vals.copy()
                v[e] += sgn * h
                js.append(c3.loss(phi_for(v, s) if which == 0 else phi_for(base, v), target))
            fd = (js[0] - js[1]) / (2 * h)
            # 1e-6 relative plus the differences' rounding noise (which binds
            # where the gradient is small against J/h)
            r = abs(fd - grad) / c3.tolerance(fd, J0, h)
            ratio = max(ratio, r)
            assert r <= 1.0, "%s entry %d: fd %.12e adjoint %.12e" % (("sigt", "s")[which], e, fd, grad)
            rel.append((abs(grad) / c3.noise(J0, h), abs(fd - grad) / max(abs(fd), 1e-300)))
            if which == 0:
                ofd = oracle_fd[e]
                n_oracle += 1
                ro = abs(ofd - grad) / c3.tolerance(ofd, J0, h)
                ratio_oracle = max(ratio_oracle, ro)
                assert ro <= 1.0, "sigt entry %d: oracle fd %.12e adjoint %.12e" % (e, ofd, grad)
        rel.sort(reverse=True)
        _say("64 parameters: max |fd - adjoint| / (1e-6 |fd| + noise) %.3f against the library's forward run, "
             "%.3f against the oracle's (%d sigma_t points); relative error at the 8 gradients furthest above the "
             "noise: %s" % (ratio, ratio_oracle, n_oracle, " ".join("%.1e" % r for _, r in rel[:8])))
        assert n_oracle == 32
        for b in bufs + [wb]:
            b.free()
    finally:
        sv.close()


# --------------------------------------------------------------------------
# differentiable, exiting-flux loss, two fields per call
# --------------------------------------------------------------------------
C4 = dict(zones=(48, 48, 48), groups=64, quad=(8, 16), legendre=3, layout="DGZ", dset=8, gset=1, zset=(1, 1, 1),
          geometry=1, seed=20261203)


@pytest.mark.timeout(600)
def test_c4_exit_flux_gradient(ctx):
    """Field 0 (the truth) is sigma_t scaled by a seeded factor in [0.98,
    1.02] on top of field 1's (the candidate) factor in [0.8, 1.2], sigma_s =
    0.5 sigma_t in both.  J = sum of (candidate - truth)^2 over the exiting
    i-faces of every octant inside an 8x8 window of (j, k): a loss of the
    workload's kind whose parameters near the window move it well above the
    rounding noise of a central difference (over all six faces J sums 1.8e7
    residuals and a single zone's share sinks into that noise).  The 8
    parameters (sigma_t and s alternating) lie in the zones next to the two
    i-faces inside the window, with sigma_t >= 0.05 /cm."""
    sv = DifferentiableSolver(ctx, ka.input_vars(**C4), 20, 4, fields=2)
    try:
        n = 48
        base = chunk_get(ctx, sv.param_ptrs("sigt")[0], sv.nparam[0])
        rng = np.random.default_rng(20261205)
        cand = base * (1.0 + 0.2 * (2.0 * rng.random(base.size) - 1.0))
        truth = cand * (1.0 + 0.02 * (2.0 * rng.random(base.size) - 1.0))
        bt = [DeviceBuffer(ctx, base.size) for _ in range(2)]
        bs = [DeviceBuffer(ctx, base.size) for _ in range(2)]
        for f, v in enumerate((truth, cand)):
            chunk_set(ctx, bt[f].ptr.value, v)
            chunk_set(ctx, bs[f].ptr.value, 0.5 * v)
            sv.set_params([bt[f].ptr.value], [bs[f].ptr.value], field=f)
        ex = sv.exits
        # DGZ i-plane chunk: [d][g][k*nj + j]; window j, k in [20, 28)
        win = np.zeros((n, n), dtype=bool)
        win[20:28, 20:28] = True
        masks = [np.broadcast_to(win.ravel()[None, :], (cnt // (n * n), n * n)).ravel() if name == "i_plane"
                 else np.zeros(cnt, dtype=bool) for _, name, cnt in ex]
        faces = [DeviceBuffer(ctx, cnt) for _ in range(2) for _, _, cnt in ex]
        sv.forward_fields(0, 2, None, [b.ptr.value for b in faces])
        target = [chunk_get(ctx, b.ptr.value, b.n) for b in faces[:len(ex)]]
        got = [chunk_get(ctx, b.ptr.value, b.n) for b in faces[len(ex):]]

        def loss(vals):
            return sum(float(np.dot((v - t)[m], (v - t)[m])) for v, t, m in zip(vals, target, masks))

        J0 = loss(got)
        fb = [DeviceBuffer(ctx, cnt) for _, _, cnt in ex]
        for b, v, t, m in zip(fb, got, target, masks):
            chunk_set(ctx, b.ptr.value, np.where(m, 2.0 * (v - t), 0.0))
        gt, gs = DeviceBuffer(ctx, base.size), DeviceBuffer(ctx, base.size)
        sv.backward_fields(1, 1, None, [b.ptr.value for b in fb], [gt.ptr.value], [gs.ptr.value])
        grads = (chunk_get(ctx, gt.ptr.value, base.size), chunk_get(ctx, gs.ptr.value, base.size))
        nz = n ** 3
        prng = np.random.default_rng(20261206)
        pb = [DeviceBuffer(ctx, base.size), DeviceBuffer(ctx, base.size)]
        params = []
        while len(params) < 8:
            i = int(prng.choice([0, 1, n - 2, n - 1]))
            j, k = (int(v) for v in prng.integers(20, 28, size=2))
            e = int(prng.integers(0, C4["groups"])) * nz + (k * n + j) * n + i
            if cand[e] >= 0.05:
                params.append((len(params) % 2, e))
        ratio, clear = 0.0, 0
        for which, e in params:
            vals = (cand, 0.5 * cand)[which]
            h = 1e-6 * abs(vals[e])
            js = []
            for sgn in (1.0, -1.0):
                v = vals.copy()
                v[e] += sgn * h
                st, ss = (v, 0.5 * cand) if which == 0 else (cand, v)
                chunk_set(ctx, pb[0].ptr.value, st)
                chunk_set(ctx, pb[1].ptr.value, ss)
                sv.set_params([pb[0].ptr.value], [pb[1].ptr.value], field=1)
                sv.forward(1)
                js.append(loss([chunk_get(ctx, p_, cnt) for _, _, p_, cnt in sv.exit_faces()]))
            fd = (js[0] - js[1]) / (2 * h)
            g = grads[which][e]
            noise = c3.noise(J0, h)
            r = abs(fd - g) / (1e-6 * abs(fd) + noise)
            ratio = max(ratio, r)
            clear += abs(g) > 100 * noise
            _say("%s entry %d: fd %.12e adjoint %.12e (relative %.1e, gradient / noise %.1e)" % (
                ("sigt", "s")[which], e, fd, g, abs(fd - g) / max(abs(fd), 1e-300), abs(g) / noise))
            assert r <= 1.0
        _say("J0 %.6e; max |fd - adjoint| / (1e-6 |fd| + noise) %.3f; %d of 8 gradients 100x above the noise"
             % (J0, ratio, clear))
        assert clear >= 6
        for b in bt + bs + faces + fb + [gt, gs] + pb:
            b.free()
    finally:
        sv.close()
