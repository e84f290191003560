"""Benchmark: Kripke source iteration on one MI355X.

A step is one source iteration (SteadyStateSolver.cpp:44-104) of the
on-device path: LTimes, scattering, source, LPlusTimes, the zone-set
pipelined sweep and the population, over the workload

    64^3 zones (8 zone sets of 32^3), 128 groups, 128 directions (Gauss-
    Legendre 8 polar x 16 azimuthal, 8 direction sets), P4 scattering
    (25 moments), FP64, DGZ nesting, seeded 3-material box geometry.

Further legs measure the metric's adjoint/forward cost on the
differentiable solve (run_adjoint) and the batched persistent launch of 48
small problems (run_batched).  Prints one JSON line.  `value` is the
reference figure of merit
(src/kripke.cpp:495-505): unknowns (D*G*Z) per second per iteration, with all
fields resident in HBM when the timed region starts.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "sweep unknowns/s FP64 (64^3z x128g x128d), % HBM peak; adjoint/forward cost"
HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s
SWEEP_STAGES = 4   # zone-set dependency levels of a 2x2x2 decomposition: one sweep dispatch each
NESTINGS = ["DGZ", "DZG", "GDZ", "GZD", "ZDG", "ZGD"]
NESTING_STEPS, NESTING_WARMUP = 5, 2  # the other nestings of the workload, each timed on its own

WORKLOAD = dict(zones=(64, 64, 64), groups=128, quad=(8, 16), legendre=4, dset=8, gset=1, zset=(2, 2, 2),
                geometry=1, seed=20261202)
# bounded sample of the same workload for the CPU oracle (about 10-30 s of CPU work)
CPU_SAMPLE = dict(zones=(32, 32, 32), groups=64, quad=(8, 16), legendre=4, dset=8, gset=1, zset=(2, 2, 2),
                  geometry=1, seed=20261202)


def algorithmic_bytes(nd_total, ng, nz, moments):
    """Compulsory HBM bytes per iteration of the metric path at the kept
    entry-point boundary: LTimes reads psi (8N) and writes phi (8P),
    LPlusTimes reads phi_out (8P) and writes rhs (8N), the sweep reads rhs
    and writes psi (16N) and reads sigma_t once (8GZ): 32N + 16P + 8GZ."""
    N = nd_total * ng * nz
    P = moments * ng * nz
    return {"ltimes": 8 * N + 8 * P, "lplustimes": 8 * P + 8 * N, "sweep": 16 * N + 8 * ng * nz, "N": N, "P": P}


def measured_traffic(layout):
    """HBM bytes per k_sweep_rows dispatch from the committed rocprofv3
    FETCH_SIZE/WRITE_SIZE summary (copy-kernel calibrated), for the nesting it
    was measured on; None otherwise."""
    path = os.path.join(ROOT, "profiles", "k_sweep_rows_%s_c1_hbm_traffic.txt" % layout.lower())
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch k_sweep_rows:"):
                    return int(line.split(":")[1])
    except OSError:
        pass
    return None


def run_gpu(layout, steps, warmup):
    """Times `steps` source iterations of the workload in `layout` after
    `warmup` untimed ones; then a breakdown pass with the hipEvent timers,
    outside the timed region.  Setup (outside the timed region) builds the
    problem and runs the library's sweep autotuner on it
    (kr_problem_autotune: kernel, patch and work-unit size)."""
    import kripke_amd as ka
    ctx = ka.Context(0)
    t0 = time.time()
    hp = ka.HostProblem(ka.input_vars(layout=layout, **WORKLOAD), ctx)
    k = hp.kernels()
    # one iteration first, so that the autotuner times its candidates on the
    # workload's own fluxes rather than on zeros
    hp.init_psi()
    hp.iterate(1)
    tune = k.autotune()
    setup_s = time.time() - t0
    hp.init_psi()
    for _ in range(warmup):
        hp.iterate(1)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        hp.iterate(1)  # ends with the population read-back (a host sync per iteration)
    ctx.synchronize()
    elapsed = time.perf_counter() - t0
    k.set_timing(True)
    k.timer_reset()
    nb = 2
    for _ in range(nb):
        hp.iterate(1)
    names = ["LTimes", "Scattering", "Source", "LPlusTimes", "SweepSolver", "k_sweep", "Population"]
    timers = {n: k.timer(n) for n in names}
    k.set_timing(False)
    hp.close()
    ctx.close()
    return elapsed, setup_s, timers, nb, tune


SWEEP_KERNELS = {0: "k_sweep_pop", 1: "k_sweep_rows", 3: "k_sweep_waves", 4: "k_sweep_pop (k-ring)"}


def nesting_figures(layout, steps, elapsed, timers, nb, tune):
    """Whole-iteration FOM, the metric path (sweep + LTimes + LPlusTimes)
    against its algorithmic bytes, and the sweep kernel's own fraction."""
    nz = 64 ** 3
    ab = algorithmic_bytes(128, 128, nz, 25)
    sweep_t, sweep_n = timers["k_sweep"]
    lt_t, _ = timers["LTimes"]
    lp_t, _ = timers["LPlusTimes"]
    ss_t, _ = timers["SweepSolver"]
    path_t = (lt_t + lp_t + ss_t) / nb
    path_bytes = ab["ltimes"] + ab["lplustimes"] + ab["sweep"]
    return {
        "unknowns_per_s": ab["N"] * steps / elapsed, "ms_per_iteration": elapsed / steps * 1e3,
        "metric_path": {"ms_per_iteration": path_t * 1e3, "alg_bytes_per_iteration": path_bytes,
                        "hbm_frac": path_bytes / path_t / HBM_PEAK if path_t > 0 else None,
                        "unknowns_per_s": ab["N"] / path_t if path_t > 0 else None},
        "sweep_kernel": {"kernel": SWEEP_KERNELS.get(tune[0], str(tune[0])),
                         "autotune_choice": list(tune),
                         "avg_launch_ms": sweep_t / max(sweep_n, 1) * 1e3,
                         "hbm_frac": (ab["sweep"] * nb) / sweep_t / HBM_PEAK if sweep_t > 0 else None},
        "breakdown_ms": {n: timers[n][0] / nb * 1e3 for n in timers},
    }


ADJ_WORKLOAD = dict(zones=(32, 32, 32), groups=32, quad=(8, 12), legendre=2, dset=8, gset=1, zset=(1, 1, 1),
                    geometry=1, seed=20261202)
ADJ_ITERATIONS, ADJ_CHECKPOINT = 10, 2


def run_adjoint(layout):
    """Adjoint/forward cost of the differentiable solve: 10 source iterations
    of 32^3 zones x 32 groups x 96 directions (GL 8x12), P2, as a function of
    per-zone sigma_t and sigma_s (one value per zone and group), loss =
    sum over a seeded 4^3 detector box of (phi - phi_target)^2; phi
    checkpoints every 2 iterations.  forward = the 10 iterations (with the
    checkpoint copies); backward = the reverse sweep including the
    recomputation from checkpoints."""
    import numpy as np
    import kripke_amd as ka
    from kripke_amd.adjoint import DeviceBuffer, DifferentiableSolver, chunk_get, chunk_set
    ctx = ka.Context(0)
    sv = DifferentiableSolver(ctx, ka.input_vars(layout=layout, **ADJ_WORKLOAD), ADJ_ITERATIONS, ADJ_CHECKPOINT)
    rng = np.random.default_rng(20261202)
    sigt = [chunk_get(ctx, p, n) for p, n in zip(sv.param_ptrs("sigt"), sv.nparam)]
    bt = [DeviceBuffer(ctx, n) for n in sv.nparam]
    bs = [DeviceBuffer(ctx, n) for n in sv.nparam]
    for b, v in zip(bt, sigt):
        chunk_set(ctx, b.ptr.value, v)
    for b, v in zip(bs, sigt):
        chunk_set(ctx, b.ptr.value, 0.5 * v)
    sv.set_params([b.ptr.value for b in bt], [b.ptr.value for b in bs])
    # detector box: 4^3 zones at a seeded corner; target phi = the phi of a solve with sigma_t perturbed by a
    # seeded +-20 % per zone and group
    nx = ADJ_WORKLOAD["zones"][0]
    c0 = rng.integers(0, nx - 4, size=3)
    bp = [DeviceBuffer(ctx, n) for n in sv.nparam]
    for b, v in zip(bp, sigt):
        chunk_set(ctx, b.ptr.value, v * (1.0 + 0.2 * (2.0 * rng.random(v.size) - 1.0)))
    sv.set_params([b.ptr.value for b in bp], [b.ptr.value for b in bs])
    sv.forward()
    target = [chunk_get(ctx, p, n) for p, n in zip(sv.phi_ptrs(), sv.nphi)]
    sv.set_params([b.ptr.value for b in bt], [b.ptr.value for b in bs])
    M, G = 9, ADJ_WORKLOAD["groups"]
    mask = np.zeros((nx, nx, nx), dtype=bool)
    mask[c0[0]:c0[0] + 4, c0[1]:c0[1] + 4, c0[2]:c0[2] + 4] = True
    zmask = np.transpose(mask, (2, 1, 0)).ravel()  # zone index z = (k*ny + j)*nx + i
    axes = {"DGZ": "agz", "DZG": "azg", "GDZ": "gaz", "GZD": "gza", "ZDG": "zag", "ZGD": "zga"}[layout]
    shape = [{"a": M, "g": G, "z": zmask.size}[ch] for ch in axes]
    zm = np.broadcast_to(zmask.reshape([zmask.size if ch == "z" else 1 for ch in axes]), shape).ravel()
    wb = [DeviceBuffer(ctx, n) for n in sv.nphi]
    times = {"forward": [], "backward": []}
    for rep in range(3):
        ctx.synchronize()
        t0 = time.perf_counter()
        sv.forward()
        ctx.synchronize()
        times["forward"].append(time.perf_counter() - t0)
        phi = [chunk_get(ctx, p, n) for p, n in zip(sv.phi_ptrs(), sv.nphi)]
        loss = sum(float(np.sum(((v - tg) ** 2)[zm])) for v, tg in zip(phi, target))
        for b, v, tg in zip(wb, phi, target):
            chunk_set(ctx, b.ptr.value, np.where(zm, 2.0 * (v - tg), 0.0))
        ctx.synchronize()
        t0 = time.perf_counter()
        sv.backward([b.ptr.value for b in wb])
        ctx.synchronize()
        times["backward"].append(time.perf_counter() - t0)
    g = chunk_get(ctx, sv.param_ptrs("grad_sigt")[0], sv.nparam[0])
    for b in bt + bs + wb + bp:
        b.free()
    sv.close()
    ctx.close()
    tf, tb = min(times["forward"]), min(times["backward"])
    return {"workload": "32^3 zones, 32 groups, 96 directions GL 8x12, P2 (9 moments), %s, 10 iterations from "
                        "psi=0, phi checkpoints every 2; parameters: sigma_t and sigma_s per zone and group; loss on "
                        "a seeded 4^3 detector box against phi of a seeded +-20%% sigma_t perturbation" % layout,
            "forward_ms": tf * 1e3, "backward_ms": tb * 1e3, "adjoint_forward_cost": tb / tf,
            "loss": loss, "grad_sigt_norm": float(np.linalg.norm(g))}


ADJB_WORKLOAD = dict(zones=(48, 48, 48), groups=64, quad=(8, 16), legendre=3, dset=8, gset=1, zset=(1, 1, 1),
                     geometry=1, seed=20261203)
ADJB_FIELDS, ADJB_ITERATIONS, ADJB_CHECKPOINT = 8, 20, 4


def run_adjoint_batched(layout):
    """Gradient throughput over candidate cross-section fields: 8 fields of
    48^3 zones x 64 groups x 128 directions (GL 8x16), P3, 20 source
    iterations from psi = 0 with phi checkpoints every 4.  Field f has
    sigma_t(g,z) x (1 + 0.2 u_f) (u_f seeded uniform in [-1, 1]) and
    sigma_s = 0.5 sigma_t; the loss is the squared mismatch of the angular flux
    leaving the domain through all 6 faces against the exiting flux of a
    seeded 'truth' field.  All 8 fields go through one kr_diff_forward call
    and one kr_diff_backward call (each field keeps its own parameters and
    phi checkpoints; the fields take turns on one pair of problems, since all
    8 live at once would hold about 300 GB)."""
    import numpy as np
    import kripke_amd as ka
    from kripke_amd.adjoint import DeviceBuffer, DifferentiableSolver, chunk_get, chunk_set
    ctx = ka.Context(0)
    F = ADJB_FIELDS
    sv = DifferentiableSolver(ctx, ka.input_vars(layout=layout, **ADJB_WORKLOAD), ADJB_ITERATIONS, ADJB_CHECKPOINT,
                              fields=F + 1)  # field F: the truth field
    sv.kf.autotune()
    sv.ka.autotune()
    rng = np.random.default_rng(20261203)
    sigt = [chunk_get(ctx, p, n) for p, n in zip(sv.param_ptrs("sigt"), sv.nparam)]
    bt = [DeviceBuffer(ctx, n) for n in sv.nparam]
    bs = [DeviceBuffer(ctx, n) for n in sv.nparam]

    def set_field(f, scale):
        for b, c, v in zip(bt, bs, sigt):
            pert = v * scale[:v.size]
            chunk_set(ctx, b.ptr.value, pert)
            chunk_set(ctx, c.ptr.value, 0.5 * pert)
        sv.set_params([b.ptr.value for b in bt], [b.ptr.value for b in bs], field=f)

    nparam = max(sv.nparam)
    set_field(F, 1.0 + 0.2 * (2.0 * rng.random(nparam) - 1.0))  # the truth field
    for f in range(F):
        set_field(f, 1.0 + 0.2 * (2.0 * rng.random(nparam) - 1.0))
    exits = sv.exit_faces()
    ne = len(exits)
    sv.forward(F)
    target = [chunk_get(ctx, p, n) for _, _, p, n in exits]
    faces = [DeviceBuffer(ctx, n) for _ in range(F) for _, _, _, n in exits]   # field-major exiting faces
    fbar = [DeviceBuffer(ctx, n) for _ in range(F) for _, _, _, n in exits]    # their cotangents
    gts = [DeviceBuffer(ctx, n) for _ in range(F) for n in sv.nparam]
    gss = [DeviceBuffer(ctx, n) for _ in range(F) for n in sv.nparam]
    tf = tb = 0.0
    losses = []
    for rep in range(2):  # the first pass is a warm-up
        ctx.synchronize()
        t0 = time.perf_counter()
        sv.forward_fields(0, F, None, [b.ptr.value for b in faces])
        ctx.synchronize()
        t1 = time.perf_counter()
        losses = []
        for f in range(F):
            loss = 0.0
            for e in range(ne):
                b = faces[f * ne + e]
                dlt = chunk_get(ctx, b.ptr.value, b.n) - target[e]
                loss += float(np.dot(dlt, dlt))
                chunk_set(ctx, fbar[f * ne + e].ptr.value, 2.0 * dlt)
            losses.append(loss)
        ctx.synchronize()
        t2 = time.perf_counter()
        sv.backward_fields(0, F, None, [b.ptr.value for b in fbar], [b.ptr.value for b in gts],
                           [b.ptr.value for b in gss])
        ctx.synchronize()
        t3 = time.perf_counter()
        tf, tb, th = t1 - t0, t3 - t2, t2 - t1
    g = chunk_get(ctx, gts[(F - 1) * sv.npr].ptr.value, sv.nparam[0])
    for b in bt + bs + faces + fbar + gts + gss:
        b.free()
    sv.close()
    ctx.close()
    return {"workload": "%d fields x (48^3 zones, 64 groups, 128 directions GL 8x16, P3 (16 moments), %s), %d "
                        "iterations from psi=0, phi checkpoints every %d; parameters sigma_t and sigma_s per zone "
                        "and group; loss on the exiting angular flux of all 6 boundary faces; all fields in one "
                        "kr_diff_forward and one kr_diff_backward call" % (F, layout, ADJB_ITERATIONS, ADJB_CHECKPOINT),
            "fields_per_s": F / (tf + tb), "fields_per_s_wall": F / (tf + tb + th),
            "forward_ms_per_field": tf / F * 1e3, "backward_ms_per_field": tb / F * 1e3,
            "adjoint_forward_cost": tb / tf, "losses": losses, "grad_sigt_norm_last": float(np.linalg.norm(g)),
            "note": "fields_per_s counts the forward and backward calls; fields_per_s_wall adds the host-side "
                    "loss and cotangent (exiting faces copied to the host)"}


BATCH_PROBLEMS, BATCH_ITERATIONS = 48, 15


def batch_configs(n=BATCH_PROBLEMS, base_seed=20261202):
    """The batched workload: n independent problems, per-problem seed
    base + index; zones per axis n0 + jitter with n0 drawn from 8..40 and a
    per-axis jitter of -2..2 (clipped to 8..40), G in {8,16,32,64}, D in
    {32,64,96,128} (GL, 8 polar angles), Legendre order 1..4.  The nesting is
    left to the autotuner (run_batched)."""
    import numpy as np
    rng = np.random.default_rng(base_seed)
    out = []
    for i in range(n):
        n0 = int(rng.integers(8, 41))
        zones = tuple(int(v) for v in np.clip(n0 + rng.integers(-2, 3, size=3), 8, 40))
        out.append(dict(zones=zones, groups=int(rng.choice([8, 16, 32, 64])),
                        quad=(8, int(rng.choice([4, 8, 12, 16]))), legendre=int(rng.integers(1, 5)), layout="DGZ",
                        dset=8, gset=1, zset=(1, 1, 1), geometry=1, seed=base_seed + i))
    return out


def run_batched():
    """48 independent problems x 15 iterations: one hipGraph launch (each
    problem's nesting and sweep shape picked by the autotuner) against the
    same iterations issued problem by problem, one call at a time."""
    import numpy as np
    import kripke_amd as ka
    ctx = ka.Context(0)
    cfgs = batch_configs()
    t0 = time.perf_counter()
    for c in cfgs:
        c["layout"] = ka.autotune_nesting(ctx, ka.input_vars(**c))[0]
    nest_s = time.perf_counter() - t0
    hps = [ka.HostProblem(ka.input_vars(**c), ctx) for c in cfgs]
    t0 = time.perf_counter()
    for hp in hps:
        hp.init_psi()
        hp.iterate(1)  # the autotuner then times its candidates on nonzero fluxes
        hp.kernels().autotune()
    tune_s = time.perf_counter() - t0
    unknowns = sum(c["zones"][0] * c["zones"][1] * c["zones"][2] * c["groups"] * c["quad"][0] * c["quad"][1]
                   for c in cfgs)
    # compulsory HBM bytes of one iteration of every problem (the metric path's 32N + 16P + 8GZ, each problem's
    # own N, P and G*Z): the graph's and the per-problem leg's fractions of the HBM peak
    alg_bytes = 0
    for c in cfgs:
        nzc = c["zones"][0] * c["zones"][1] * c["zones"][2]
        b = algorithmic_bytes(c["quad"][0] * c["quad"][1], c["groups"], nzc, (c["legendre"] + 1) ** 2)
        alg_bytes += b["ltimes"] + b["lplustimes"] + b["sweep"]
    # problem by problem
    for hp in hps:
        hp.init_psi()
    ctx.synchronize()
    t0 = time.perf_counter()
    seq = np.array([[hp.iterate(1) for _ in range(BATCH_ITERATIONS)] for hp in hps])
    seq_s = time.perf_counter() - t0
    # one graph launch (the first call captures; time a replay from psi = 0)
    b = ka.Batch(ctx, hps)
    for hp in hps:
        hp.init_psi()
    b.iterate(BATCH_ITERATIONS)
    for hp in hps:
        hp.init_psi()
    ctx.synchronize()
    t0 = time.perf_counter()
    pops = b.iterate(BATCH_ITERATIONS)
    graph_s = time.perf_counter() - t0
    same = bool(np.array_equal(pops, seq))
    b.close()
    for hp in hps:
        hp.close()
    ctx.close()
    nest = {l: sum(1 for c in cfgs if c["layout"] == l) for l in ka.LAYOUTS}
    return {"workload": "%d problems (zones n0+-2 per axis, n0 in 8..40; G in {8,16,32,64}, D in {32,64,96,128}, "
                        "P1-P4; nesting and sweep shape autotuned), %d iterations each" % (BATCH_PROBLEMS,
                                                                                      BATCH_ITERATIONS),
            "unknowns_per_iteration": unknowns, "graph_ms": graph_s * 1e3, "per_problem_ms": seq_s * 1e3,
            "speedup_vs_per_problem": seq_s / graph_s, "graph_unknowns_per_s": unknowns * BATCH_ITERATIONS / graph_s,
            "alg_bytes_per_iteration": alg_bytes,
            "graph_hbm_frac": alg_bytes * BATCH_ITERATIONS / graph_s / HBM_PEAK,
            "per_problem_hbm_frac": alg_bytes * BATCH_ITERATIONS / seq_s / HBM_PEAK,
            "populations_equal_per_problem_runs": same, "nestings": {k: v for k, v in nest.items() if v},
            "autotune_nesting_s": nest_s, "autotune_s": tune_s}


def run_cpu_baseline():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    o = po.Oracle(po.make_input(layout="DGZ", **CPU_SAMPLE), kind="ref", omp=True)
    o.init_psi()
    iters = 2
    o.timer_reset()
    for _ in range(iters):
        o.iterate()
    solve, _ = o.timer("Solve")
    nz = CPU_SAMPLE["zones"][0] * CPU_SAMPLE["zones"][1] * CPU_SAMPLE["zones"][2]
    N = nz * CPU_SAMPLE["groups"] * CPU_SAMPLE["quad"][0] * CPU_SAMPLE["quad"][1]
    path = sum(o.timer(n)[0] for n in ("LTimes", "LPlusTimes", "SweepSolver"))
    return {"value": N * iters / solve, "unit": "unknowns/s", "cores": po.omp_threads(), "kind": "port",
            "sample": "CPU oracle (OpenMP restatement of the reference OpenMP loop orders), 32^3 zones x 64 groups "
                      "x 128 directions, P4, 8 zone sets, DGZ, %d iterations; whole-Solve FOM" % iters,
            "metric_path_unknowns_per_s": N * iters / path, "solve_s": solve}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--layout", default="DGZ")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-nestings", action="store_true", help="skip the other five nestings of the workload")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint/forward cost leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-launch leg")
    ap.add_argument("--batch-only", action="store_true",
                    help="run only the batched-launch leg and print its record (for profiling that leg)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one MI355X: this build has no multi-GPU path (use --gpus 1)", file=sys.stderr)
        sys.exit(2)
    if int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs as a single process on one GPU", file=sys.stderr)
        sys.exit(2)

    if args.batch_only:
        print(json.dumps({"batched": run_batched()}))
        return

    elapsed, setup_s, timers, nb, tune = run_gpu(args.layout, args.steps, args.warmup)
    fig = nesting_figures(args.layout, args.steps, elapsed, timers, nb, tune)
    ab = algorithmic_bytes(128, 128, 64 ** 3, 25)
    N = ab["N"]
    sweep_t, sweep_n = timers["k_sweep"]
    # sweep kernel roofline: algorithmic bytes of all stage launches / their time
    sweep_achieved = (ab["sweep"] * nb) / sweep_t if sweep_t > 0 else None
    out = {
        "metric": METRIC,
        "value": fig["unknowns_per_s"],
        "unit": "unknowns/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": fig["ms_per_iteration"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded 3-material boxes in a 100 cm cube, vacuum boundaries)",
        "config": {"workload": "64^3 zones (2x2x2 zone sets), 128 groups, 128 directions GL 8x16 (8 direction sets), "
                               "P4 (25 moments), %s nesting, on-device source iteration, sweep shape autotuned"
                               % args.layout,
                   "unknowns": N, "nesting": args.layout, "zone_sets": [2, 2, 2], "parallelism": "1 GPU"},
        "roofline": {"bound": "hbm", "achieved": sweep_achieved / 1e9 if sweep_achieved else None,
                     "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": sweep_achieved / HBM_PEAK if sweep_achieved else None,
                     "traffic": measured_traffic(args.layout),
                     "kernel": fig["sweep_kernel"]["kernel"],
                     "autotune_choice": fig["sweep_kernel"]["autotune_choice"],
                     "alg_bytes_per_launch": ab["sweep"] / SWEEP_STAGES,
                     "avg_launch_ms": sweep_t / max(sweep_n, 1) * 1e3},
        "metric_path": fig["metric_path"],
        "breakdown_ms": fig["breakdown_ms"],
        "setup_s": setup_s,
    }
    if not args.no_nestings:
        # the same workload in every nesting (the headline's own figures included once more, from its run above)
        per = {args.layout: fig}
        for lay in NESTINGS:
            if lay != args.layout:
                e, _, tm, b, tu = run_gpu(lay, NESTING_STEPS, NESTING_WARMUP)
                per[lay] = nesting_figures(lay, NESTING_STEPS, e, tm, b, tu)
        out["nestings"] = {"steps": NESTING_STEPS, "warmup": NESTING_WARMUP,
                           "figures": {lay: per[lay] for lay in NESTINGS if lay in per}}
    if not args.no_adjoint:
        out["adjoint"] = run_adjoint(args.layout)
        out["adjoint_batched"] = run_adjoint_batched(args.layout)
    if not args.no_batch:
        out["batched"] = run_batched()
    if not args.no_cpu:
        out["cpu_baseline"] = run_cpu_baseline()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
