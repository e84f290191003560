"""The kripke command (src/kripke.cpp): option parsing and validation on the
CPU; on the GPU, the printed per-iteration particle counts against the
oracle and the report sections (Timers, TIMER_NAMES/TIMER_DATA, Figures of
Merit, END)."""
import os
import re
import subprocess

import pytest

import pyoracle as po

HERE = os.path.dirname(os.path.abspath(__file__))
KRIPKE = os.path.join(os.path.dirname(HERE), "kripke_amd", "kripke")


def run(*args, timeout=120):
    if not os.path.exists(KRIPKE):
        pytest.fail("kripke command not built (run __graft_entry__.build())")
    r = subprocess.run([KRIPKE] + [str(a) for a in args], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True, timeout=timeout)
    return r.returncode, r.stdout


def test_usage():
    rc, out = run("--help")
    assert rc == 1
    for opt in ("--groups", "--legendre", "--quad", "--zones", "--sigt", "--sigs", "--arch", "--layout", "--procs",
                "--dset", "--gset", "--zset", "--niter", "--pmethod"):
        assert opt in out, opt


@pytest.mark.parametrize("args,msg", [
    (["--groups", "5", "--gset", "2"], "Number of groups (5) must be evenly divided by number of groupsets (2)"),
    (["--quad", "4"], "Number of directions (4) needs to be at least 8"),
    (["--dset", "3"], "Number of direction sets (3) must be a multiple of 8"),
    (["--legendre", "-1"], "Legendre scattering order (-1) must be >= 0"),
    (["--niter", "0"], "You must run at least one iteration (0)"),
    (["--zset", "0,1,1"], "Number of zone-sets in each dim need to be greater than or equal to 1"),
    (["--procs", "2,1,1"], "--procs must be 1,1,1"),
    (["--layout", "XYZ"], "Unknown layout XYZ"),
    (["--arch", "Sequential"], "not built here"),
    (["--bogus"], "Unknown option --bogus"),
])
def test_bad_arguments(args, msg):
    rc, out = run(*args)
    assert rc == 1
    assert msg in out


def _counts(out):
    return [float(m.group(1)) for m in re.finditer(r"iter \d+: particle count=(\S+), change=", out)]


CFG = dict(zones=(8, 8, 8), groups=8, quad=(4, 4), legendre=2, dset=8, gset=2, zset=(2, 1, 1), geometry=1,
           seed=20261202)
ARGS = ["--zones", "8,8,8", "--groups", 8, "--quad", "4:4", "--legendre", 2, "--dset", 8, "--gset", 2, "--zset",
        "2,1,1", "--geometry", "boxes", "--seed", 20261202, "--niter", 3]


def _oracle_counts(n, **kw):
    o = po.Oracle(po.make_input(**kw), kind="ref", omp=True)
    o.init_psi()
    return [o.iterate()[1] for _ in range(n)]


@pytest.mark.gpu
@pytest.mark.parametrize("layout,mode,pmethod", [("DGZ", "pipeline", "sweep"), ("GZD", "ref", "sweep"),
                                                 ("ZDG", "pipeline", "bj"), ("DZG", "ref", "bj")])
def test_counts_match_oracle(layout, mode, pmethod):
    rc, out = run(*(ARGS + ["--layout", layout, "--mode", mode, "--pmethod", pmethod]))
    assert rc == 0, out[-2000:]
    got = _counts(out)
    want = _oracle_counts(3, layout=layout, pmethod=pmethod, **CFG)
    assert len(got) == 3
    for g, w in zip(got, want):
        assert abs(g - w) <= 1e-6 * abs(w), (g, w)  # printed with %e (7 significant digits)
    assert ("Block solve method:    %s" % ("Block Jacobi" if pmethod == "bj" else "Sweep")) in out
    names = re.search(r"TIMER_NAMES:(\S+)", out).group(1).split(",")
    data = re.search(r"TIMER_DATA:(\S+)", out).group(1).split(",")
    assert names == sorted(names) and len(names) == len(data)
    for t in ("Generate", "LPlusTimes", "LTimes", "Scattering", "Solve", "Source", "SweepSolver"):
        assert t in names, t
    if mode == "ref":
        assert "SweepSubdomain" in names and "Population" in names
    assert re.search(r"SweepSolver\s+3\s", out)
    fom = re.search(r"Throughput:\s+(\S+) \[unknowns/\(second/iteration\)\]", out)
    assert fom and float(fom.group(1)) > 0
    eff = re.search(r"Sweep efficiency :\s+(\S+) \[", out)
    if "SweepSubdomain" in names:
        assert 0.0 < float(eff.group(1)) <= 100.0, eff.group(0)
    else:  # the pipeline makes no per-subdomain sweep calls
        assert eff.group(1) == "n/a", eff.group(0)
    assert "Number of unknowns: %d" % (8 * 16 * 8 ** 3) in out
    assert out.rstrip().endswith("END")


@pytest.mark.gpu
def test_reference_default_problem():
    """No options: the reference's default run (16^3 Kobayashi-3i, dummy S2
    with 96 directions, 32 groups in 2 sets, P4), here 2 iterations."""
    rc, out = run("--niter", 2)
    assert rc == 0, out[-2000:]
    got = _counts(out)
    want = _oracle_counts(2, zones=(16, 16, 16), groups=32, quad=None, directions=96, legendre=4, layout="DGZ",
                          dset=8, gset=2, zset=(1, 1, 1), geometry=0)
    for g, w in zip(got, want):
        assert abs(g - w) <= 1e-6 * abs(w), (g, w)
    assert "Dummy S2 with 96 points" in out


@pytest.mark.gpu
def test_layout_auto():
    """--layout auto: the autotuner's nesting is reported and the counts still
    match the oracle (the nesting changes only the memory order)."""
    rc, out = run(*(ARGS + ["--layout", "auto"]))
    assert rc == 0, out[-2000:]
    m = re.search(r"Layout chosen by the autotuner: (\w+) \((\S+) ms per source iteration\)", out)
    assert m and m.group(1) in ("DGZ", "DZG", "GDZ", "GZD", "ZDG", "ZGD") and float(m.group(2)) > 0
    assert "Data Layout:           auto" in out
    got = _counts(out)
    want = _oracle_counts(3, layout=m.group(1), pmethod="sweep", **CFG)
    for g, w in zip(got, want):
        assert abs(g - w) <= 1e-6 * abs(w), (g, w)
