"""Transposed operators and the reverse-mode gradient (GPU, through the C-ABI).

Pins, each independent of the code under test:
* transpose identities <y, A x> = <A^T y, x> on random FP64 vectors for
  LTimes, LPlusTimes, scattering (structure-aware, dense and zonal) and the
  pipelined sweep, in every nesting;
* central finite differences of the forward solve (the same library, whose
  forward is pinned to the oracle elsewhere) against the adjoint gradient;
* the gradient does not depend on the checkpoint interval (bit for bit);
* torch.autograd.gradcheck of the torch wrapper in FP64.
"""
import numpy as np
import pytest

from helpers import kwargs_for, product_for

pytestmark = pytest.mark.gpu

LAYOUTS = ["DGZ", "DZG", "GDZ", "GZD", "ZDG", "ZGD"]
TOL_DOT = 1e-13


def _rand_fields(hp, name, count, rng, offset=0.0):
    vals = []
    for c in range(count):
        v = offset + rng.standard_normal(hp.get(name, c).size)
        hp.set(name, c, v)
        vals.append(v)
    return vals


def _get(hp, name, count):
    return [hp.get(name, c) for c in range(count)]


def _dot(a, b):
    return float(sum(np.dot(x, y) for x, y in zip(a, b)))


def _scale(a, b):
    return float(sum(np.dot(np.abs(x), np.abs(y)) for x, y in zip(a, b)))


def _check_dot(lhs_pairs, rhs_pairs):
    lhs, rhs = _dot(*lhs_pairs), _dot(*rhs_pairs)
    scale = max(_scale(*lhs_pairs), 1e-300)
    assert abs(lhs - rhs) <= TOL_DOT * scale, "<y,Ax>=%.17g <A^T y,x>=%.17g (scale %.3g)" % (lhs, rhs, scale)


@pytest.mark.parametrize("layout", LAYOUTS)
def test_ltimes_transpose(ctx, layout):
    hp = product_for(ctx, **kwargs_for(layout, zset=(2, 1, 1), gset=2, groups=8))
    k = hp.kernels()
    rng = np.random.default_rng(41)
    x = _rand_fields(hp, "psi", hp.num_sdoms, rng)
    k.ltimes(accumulate=False)
    Lx = _get(hp, "phi", hp.num_pr)
    y = _rand_fields(hp, "phi", hp.num_pr, rng)
    k.ltimes_adjoint()
    Lty = _get(hp, "psi", hp.num_sdoms)
    _check_dot((y, Lx), (Lty, x))


@pytest.mark.parametrize("layout", LAYOUTS)
def test_lplustimes_transpose(ctx, layout):
    hp = product_for(ctx, **kwargs_for(layout, zset=(1, 2, 1), gset=2, groups=8))
    k = hp.kernels()
    rng = np.random.default_rng(42)
    x = _rand_fields(hp, "phi_out", hp.num_pr, rng)
    k.lplustimes(accumulate=False)
    Lx = _get(hp, "rhs", hp.num_sdoms)
    y = _rand_fields(hp, "rhs", hp.num_sdoms, rng)
    k.lplustimes_adjoint()
    Lty = _get(hp, "phi_out", hp.num_pr)
    _check_dot((y, Lx), (Lty, x))


@pytest.mark.parametrize("dense", [False, True])
@pytest.mark.parametrize("layout", ["DGZ", "DZG", "GDZ", "GZD", "ZDG", "ZGD"])
def test_scattering_transpose(ctx, layout, dense):
    hp = product_for(ctx, **kwargs_for(layout, zset=(1, 1, 2), gset=2, groups=4, legendre=1))
    k = hp.kernels()
    rng = np.random.default_rng(43)
    if dense:  # group coupling and n = 1 terms: the dense kernels
        hp.set("data/sigs", 0, rng.uniform(0.0, 0.1, hp.get("data/sigs").size))
    x = _rand_fields(hp, "phi", hp.num_pr, rng)
    k.scattering(accumulate=False)
    Sx = _get(hp, "phi_out", hp.num_pr)
    y = _rand_fields(hp, "phi_out", hp.num_pr, rng)
    k.scattering_adjoint(accumulate=False)
    Sty = _get(hp, "phi", hp.num_pr)
    _check_dot((y, Sx), (Sty, x))


def test_scattering_zonal_matches_materials(ctx):
    """s(g,z) = the mixed-element average of the generated sigma_s: the zonal
    kernel reproduces the structure-aware scattering bit for bit."""
    from kripke_amd.adjoint import DeviceBuffer, chunk_set
    hp = product_for(ctx, **kwargs_for("DGZ", zset=(2, 1, 1), groups=4))
    k = hp.kernels()
    rng = np.random.default_rng(44)
    _rand_fields(hp, "phi", hp.num_pr, rng)
    k.scattering(accumulate=False)
    ref = _get(hp, "phi_out", hp.num_pr)
    sigs = hp.get("data/sigs").reshape(3, -1, hp.inp.num_groups, hp.inp.num_groups)  # DGZ: [mat][n][g][gp]
    bufs = []
    for c in range(hp.num_pr):
        r = c  # one group set: PR chunk c is zone set c
        z2m = hp.get("zone_to_mixelem", r).astype(int)
        z2n = hp.get("zone_to_num_mixelem", r).astype(int)
        mat = hp.get("mixelem_to_material", r).astype(int)
        frac = hp.get("mixelem_to_fraction", r)
        nz, ng = z2m.size, hp.inp.num_groups
        s = np.zeros((ng, nz))
        for z in range(nz):
            acc = np.zeros(ng)
            for q in range(z2n[z]):
                m = z2m[z] + q
                acc = acc + np.diag(sigs[mat[m], 0]) * frac[m]
            s[:, z] = acc
        b = DeviceBuffer(ctx, s.size)
        chunk_set(ctx, b.ptr.value, s.ravel())
        bufs.append(b)
    k.scattering_zonal([b.ptr.value for b in bufs])
    for c in range(hp.num_pr):
        assert np.array_equal(hp.get("phi_out", c), ref[c])
    for b in bufs:
        b.free()


@pytest.mark.parametrize("layout", LAYOUTS)
@pytest.mark.parametrize("zones", [(8, 8, 8), (12, 6, 10), (128, 80, 128)])
def test_sweep_transpose(ctx, layout, zones):
    big = zones[0] > 64  # 64 x 40 x 64 subdomains: two k-patches in the row-scan sweep
    hp = product_for(ctx, **kwargs_for(layout, zset=(2, 2, 2), groups=2 if big else 8, zones=zones,
                                       quad=(4, 2) if big else (4, 4), legendre=1 if big else 2))
    k = hp.kernels()
    rng = np.random.default_rng(45)
    x = _rand_fields(hp, "rhs", hp.num_sdoms, rng)
    k.sweep_solver(list(range(hp.num_sdoms)))
    Hx = _get(hp, "psi", hp.num_sdoms)
    y = _rand_fields(hp, "psi", hp.num_sdoms, rng)
    k.sweep_solver_adjoint()
    Hty = _get(hp, "rhs", hp.num_sdoms)
    _check_dot((y, Hx), (Hty, x))


PLANES = ("i_plane", "j_plane", "k_plane")


@pytest.mark.parametrize("layout", LAYOUTS)
@pytest.mark.parametrize("zones,zset", [((12, 6, 10), (2, 2, 2)), ((8, 8, 8), (1, 1, 1)), ((128, 80, 128), (2, 2, 2)),
                                        ((1, 1, 1), (1, 1, 1)), ((7, 5, 3), (1, 1, 1)), ((1, 4, 3), (1, 2, 1)),
                                        ((10, 6, 6), (2, 2, 2))])
def test_sweep_transpose_seeded(ctx, layout, zones, zset):
    """The sweep as a map rhs -> (psi, faces leaving the domain) and its
    transpose, the seeded adjoint sweep (the faces' cotangent enters as
    adjoint inflow on the domain boundary):
    <y, H x> + <fb, B x> = <H^T y + B^T fb, x>.  Besides the tuned
    extents: a one-zone subdomain, odd rows, a single-zone axis and a 2x2x2
    pipeline of 5x3x3 subdomains."""
    big = zones[0] > 64
    hp = product_for(ctx, **kwargs_for(layout, zset=zset, groups=2 if big else 4, zones=zones,
                                       quad=(4, 2) if big else (4, 4), legendre=1))
    k = hp.kernels()
    rng = np.random.default_rng(47)
    exits = [(s, PLANES[a]) for s in range(hp.num_sdoms) for a, e in enumerate(k.sdom_exits(s)) if e]
    assert exits
    x = _rand_fields(hp, "rhs", hp.num_sdoms, rng)
    k.sweep_solver(list(range(hp.num_sdoms)))
    Hx = _get(hp, "psi", hp.num_sdoms)
    Bx = [hp.get(name, s) for s, name in exits]
    y = _rand_fields(hp, "psi", hp.num_sdoms, rng)
    fb = []
    for s, name in exits:
        v = rng.standard_normal(hp.get(name, s).size)
        hp.set(name, s, v)
        fb.append(v)
    k.sweep_solver_adjoint(seeded=True)
    Aty = _get(hp, "rhs", hp.num_sdoms)
    _check_dot((y + fb, Hx + Bx), (Aty, x))


@pytest.mark.parametrize("layout", ["DGZ", "GDZ"])
@pytest.mark.parametrize("zones,zset,quad", [((32, 24, 32), (1, 1, 2), (4, 8)), ((32, 12, 32), (1, 2, 1), (4, 8)),
                                             ((16, 16, 16), (2, 1, 1), (8, 8)), ((128, 6, 8), (1, 1, 2), (4, 4))])
def test_sweep_transpose_row_widths(ctx, layout, zones, zset, quad):
    """The adjoint sweep for rows of 16 and 8 lanes, with k-patches and zone-set
    pipelines."""
    hp = product_for(ctx, **kwargs_for(layout, zset=zset, groups=2, zones=zones, quad=quad, legendre=1))
    k = hp.kernels()
    rng = np.random.default_rng(46)
    x = _rand_fields(hp, "rhs", hp.num_sdoms, rng)
    k.sweep_solver(list(range(hp.num_sdoms)))
    Hx = _get(hp, "psi", hp.num_sdoms)
    y = _rand_fields(hp, "psi", hp.num_sdoms, rng)
    k.sweep_solver_adjoint()
    Hty = _get(hp, "rhs", hp.num_sdoms)
    _check_dot((y, Hx), (Hty, x))


# --------------------------------------------------------------------------
# gradients
# --------------------------------------------------------------------------
def _solver(ctx, layout, iterations, ck):
    import kripke_amd as ka
    from kripke_amd.adjoint import DifferentiableSolver
    inp = ka.input_vars(**kwargs_for(layout, zones=(6, 6, 6), zset=(2, 1, 1), groups=2, quad=(4, 2), legendre=1))
    return DifferentiableSolver(ctx, inp, iterations, checkpoint_every=ck)


def _params(sv, ctx):
    from kripke_amd.adjoint import DeviceBuffer, chunk_get, chunk_set
    sigt = [chunk_get(ctx, p, n) for p, n in zip(sv.param_ptrs("sigt"), sv.nparam)]
    s = [0.6 * t for t in sigt]
    bt = [DeviceBuffer(ctx, t.size) for t in sigt]
    bs = [DeviceBuffer(ctx, t.size) for t in s]
    for b, v in zip(bt + bs, sigt + s):
        chunk_set(ctx, b.ptr.value, v)
    return sigt, s, bt, bs


def _loss(sv, ctx, w):
    from kripke_amd.adjoint import chunk_get
    return sum(float(np.dot(wc, chunk_get(ctx, p, n))) for wc, p, n in zip(w, sv.phi_ptrs(), sv.nphi))


def _gradients(sv, ctx, bt, bs, w, wbufs):
    from kripke_amd.adjoint import chunk_get
    sv.set_params([b.ptr.value for b in bt], [b.ptr.value for b in bs])
    sv.forward()
    sv.backward([b.ptr.value for b in wbufs])
    gt = [chunk_get(ctx, p, n) for p, n in zip(sv.param_ptrs("grad_sigt"), sv.nparam)]
    gs = [chunk_get(ctx, p, n) for p, n in zip(sv.param_ptrs("grad_s"), sv.nparam)]
    return gt, gs


@pytest.mark.parametrize("layout", ["DGZ", "ZDG"])
def test_gradient_vs_finite_differences(ctx, layout):
    from kripke_amd.adjoint import DeviceBuffer, chunk_set
    sv = _solver(ctx, layout, iterations=3, ck=2)
    sigt, s, bt, bs = _params(sv, ctx)
    rng = np.random.default_rng(46)
    w = [rng.standard_normal(n) for n in sv.nphi]  # J = <w, phi_K>
    wbufs = [DeviceBuffer(ctx, n) for n in sv.nphi]
    for b, v in zip(wbufs, w):
        chunk_set(ctx, b.ptr.value, v)
    gt, gs = _gradients(sv, ctx, bt, bs, w, wbufs)
    sv.set_params([b.ptr.value for b in bt], [b.ptr.value for b in bs])
    sv.forward()  # the backward pass leaves intermediate states in the forward problem
    J0 = abs(_loss(sv, ctx, w))
    for which, base, grad, bufs in (("sigt", sigt, gt, bt), ("s", s, gs, bs)):
        for _ in range(6):
            c = int(rng.integers(len(base)))
            e = int(rng.integers(base[c].size))
            h = 1e-6 * abs(base[c][e])
            vals = []
            for sgn in (1.0, -1.0):
                v = base[c].copy()
                v[e] += sgn * h
                chunk_set(ctx, bufs[c].ptr.value, v)
                sv.set_params([b.ptr.value for b in bt], [b.ptr.value for b in bs])
                sv.forward()
                vals.append(_loss(sv, ctx, w))
            chunk_set(ctx, bufs[c].ptr.value, base[c])
            fd = (vals[0] - vals[1]) / (2 * h)
            # central-difference truncation ~h^2, plus the rounding of J (~1e-13 |J|) over 2h
            tol = 1e-6 * abs(fd) + 1e-13 * J0 / h
            assert abs(fd - grad[c][e]) <= tol, \
                "%s chunk %d entry %d: fd %.12e adjoint %.12e" % (which, c, e, fd, grad[c][e])
    for b in bt + bs + wbufs:
        b.free()
    sv.close()


@pytest.mark.parametrize("layout", ["DGZ", "DZG"])
def test_exiting_flux_gradient_vs_finite_differences(ctx, layout):
    """Loss on the angular flux leaving the domain after K iterations,
    J = sum over the exiting faces of (psi_exit - target)^2: the adjoint
    gradient (exiting-face cotangent as adjoint inflow of the last sweep)
    against central differences of the forward solve."""
    from kripke_amd.adjoint import DeviceBuffer, chunk_get, chunk_set
    sv = _solver(ctx, layout, iterations=3, ck=2)
    sigt, s, bt, bs = _params(sv, ctx)
    rng = np.random.default_rng(49)
    exits = sv.exit_faces()
    target = {(sd, nm): rng.uniform(0.0, 0.1, n) for sd, nm, _, n in exits}

    def loss():
        return sum(float(np.sum((chunk_get(ctx, p, n) - target[(sd, nm)]) ** 2)) for sd, nm, p, n in exits)

    sv.set_params([b.ptr.value for b in bt], [b.ptr.value for b in bs])
    sv.forward()
    J0 = loss()
    fbufs = {}
    for sd, nm, p, n in exits:
        b = DeviceBuffer(ctx, n)
        chunk_set(ctx, b.ptr.value, 2.0 * (chunk_get(ctx, p, n) - target[(sd, nm)]))
        fbufs[(sd, nm)] = b
    sv.backward(None, {key: b.ptr.value for key, b in fbufs.items()})
    gt = [chunk_get(ctx, p, n) for p, n in zip(sv.param_ptrs("grad_sigt"), sv.nparam)]
    gs = [chunk_get(ctx, p, n) for p, n in zip(sv.param_ptrs("grad_s"), sv.nparam)]
    assert any(np.any(g != 0.0) for g in gt)
    for which, base, grad, bufs in (("sigt", sigt, gt, bt), ("s", s, gs, bs)):
        for _ in range(6):
            c = int(rng.integers(len(base)))
            e = int(rng.integers(base[c].size))
            h = 1e-6 * abs(base[c][e])
            vals = []
            for sgn in (1.0, -1.0):
                v = base[c].copy()
                v[e] += sgn * h
                chunk_set(ctx, bufs[c].ptr.value, v)
                sv.set_params([b.ptr.value for b in bt], [b.ptr.value for b in bs])
                sv.forward()
                vals.append(loss())
            chunk_set(ctx, bufs[c].ptr.value, base[c])
            fd = (vals[0] - vals[1]) / (2 * h)
            tol = 1e-6 * abs(fd) + 1e-13 * J0 / h
            assert abs(fd - grad[c][e]) <= tol, \
                "%s chunk %d entry %d: fd %.12e adjoint %.12e" % (which, c, e, fd, grad[c][e])
    for b in bt + bs + list(fbufs.values()):
        b.free()
    sv.close()


def test_gradient_independent_of_checkpoint_interval(ctx):
    from kripke_amd.adjoint import DeviceBuffer, chunk_set
    ref = None
    for ck in (1, 2, 4):
        sv = _solver(ctx, "DGZ", iterations=4, ck=ck)
        sigt, s, bt, bs = _params(sv, ctx)
        w = [np.random.default_rng(48).standard_normal(n) for n in sv.nphi]
        wbufs = [DeviceBuffer(ctx, n) for n in sv.nphi]
        for b, v in zip(wbufs, w):
            chunk_set(ctx, b.ptr.value, v)
        gt, gs = _gradients(sv, ctx, bt, bs, w, wbufs)
        if ref is None:
            ref = (gt, gs)
        else:
            for a, b in zip(gt + gs, ref[0] + ref[1]):
                assert np.array_equal(a, b)
        for b in bt + bs + wbufs:
            b.free()
        sv.close()


def test_torch_gradcheck():
    """torch.autograd.gradcheck (FP64, fast mode) of the torch wrapper, in a
    child process that imports torch before the library (one HIP runtime)."""
    import os
    import subprocess
    import sys
    pytest.importorskip("torch")
    case = os.path.join(os.path.dirname(os.path.abspath(__file__)), "torch_gradcheck_case.py")
    r = subprocess.run([sys.executable, case], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=300)
    print(r.stdout[-2000:])
    assert r.returncode == 0, r.stdout[-2000:]


# --------------------------------------------------------------------------
# the C-ABI forward/backward pair (kr_diff) against the same chain sequenced
# call by call from Python
# --------------------------------------------------------------------------
def _sequenced(ctx, layout, iterations, ck):
    import kripke_amd as ka
    from adjoint_sequence import SequencedSolver
    inp = ka.input_vars(**kwargs_for(layout, zones=(6, 6, 6), zset=(2, 1, 1), groups=2, quad=(4, 2), legendre=1))
    return SequencedSolver(ctx, inp, iterations, checkpoint_every=ck)


@pytest.mark.parametrize("layout,faces", [("DGZ", False), ("GZD", False), ("DZG", True), ("ZGD", True)])
def test_diff_pair_matches_python_sequence(ctx, layout, faces):
    """kr_diff_forward/kr_diff_backward give the phi_K, exiting faces and
    gradients of the Python-sequenced chain bit for bit (phi and face
    cotangents; 5 iterations, checkpoints every 2)."""
    from kripke_amd.adjoint import DeviceBuffer, chunk_get, chunk_set
    sv = _solver(ctx, layout, iterations=5, ck=2)
    sq = _sequenced(ctx, layout, iterations=5, ck=2)
    sigt, s, bt, bs = _params(sv, ctx)
    rng = np.random.default_rng(51)
    w = [rng.standard_normal(n) for n in sv.nphi]
    wbufs = [DeviceBuffer(ctx, n) for n in sv.nphi]
    for b, v in zip(wbufs, w):
        chunk_set(ctx, b.ptr.value, v)
    out = []
    for solver in (sv, sq):
        solver.set_params([b.ptr.value for b in bt], [b.ptr.value for b in bs])
        solver.forward()
        phi = [chunk_get(ctx, p, n) for p, n in zip(solver.phi_ptrs(), solver.nphi)]
        ex = solver.exit_faces()
        fvals = [chunk_get(ctx, p, n) for _, _, p, n in ex]
        fb = None
        fbufs = []
        if faces:
            fb = {}
            for (sd, nm, _, n), v in zip(ex, fvals):
                b = DeviceBuffer(ctx, n)
                chunk_set(ctx, b.ptr.value, np.sin(v))
                fb[(sd, nm)] = b.ptr.value
                fbufs.append(b)
        solver.backward([b.ptr.value for b in wbufs], fb)
        gt = [chunk_get(ctx, p, n) for p, n in zip(solver.param_ptrs("grad_sigt"), solver.nparam)]
        gs = [chunk_get(ctx, p, n) for p, n in zip(solver.param_ptrs("grad_s"), solver.nparam)]
        out.append((phi, fvals, gt, gs, [(sd, nm) for sd, nm, _, _ in ex]))
        for b in fbufs:
            b.free()
    a, b = out
    assert a[4] == b[4], "exit lists differ"
    for x, y in zip(a[0] + a[1] + a[2] + a[3], b[0] + b[1] + b[2] + b[3]):
        assert np.array_equal(x, y)
    assert any(np.any(g != 0.0) for g in a[2] + a[3])
    for buf in bt + bs + wbufs:
        buf.free()
    sv.close()
    sq.close()


def test_diff_fields_in_one_call(ctx):
    """Three fields through one kr_diff_forward and one kr_diff_backward call
    equal the same fields run one at a time (phi_K, exiting faces,
    gradients; bit for bit); backward of a field whose parameters changed
    after its forward pass is refused."""
    import kripke_amd as ka
    from kripke_amd.adjoint import DeviceBuffer, DifferentiableSolver, chunk_get, chunk_set
    inp = ka.input_vars(**kwargs_for("DGZ", zones=(6, 6, 6), zset=(1, 2, 1), groups=2, quad=(4, 2), legendre=1))
    F = 3
    sv = DifferentiableSolver(ctx, inp, 4, checkpoint_every=3, fields=F)
    sigt = [chunk_get(ctx, p, n) for p, n in zip(sv.param_ptrs("sigt"), sv.nparam)]
    rng = np.random.default_rng(52)
    keep = []
    for f in range(F):
        bt = [DeviceBuffer(ctx, n) for n in sv.nparam]
        bs = [DeviceBuffer(ctx, n) for n in sv.nparam]
        for b, c, v in zip(bt, bs, sigt):
            pert = v * (1.0 + 0.3 * rng.uniform(-1, 1, v.size))
            chunk_set(ctx, b.ptr.value, pert)
            chunk_set(ctx, c.ptr.value, 0.4 * pert)
        sv.set_params([b.ptr.value for b in bt], [b.ptr.value for b in bs], field=f)
        keep += bt + bs
    ne, npr = len(sv.exits), sv.npr
    phik = [DeviceBuffer(ctx, n) for _ in range(F) for n in sv.nphi]
    faces = [DeviceBuffer(ctx, n) for _ in range(F) for _, _, n in sv.exits]
    wb = [DeviceBuffer(ctx, n) for _ in range(F) for n in sv.nphi]
    fb = [DeviceBuffer(ctx, n) for _ in range(F) for _, _, n in sv.exits]
    for b in wb + fb:
        chunk_set(ctx, b.ptr.value, rng.standard_normal(b.n))
    gt = [DeviceBuffer(ctx, n) for _ in range(F) for n in sv.nparam]
    gs = [DeviceBuffer(ctx, n) for _ in range(F) for n in sv.nparam]
    sv.forward_fields(0, F, [b.ptr.value for b in phik], [b.ptr.value for b in faces])
    sv.backward_fields(0, F, [b.ptr.value for b in wb], [b.ptr.value for b in fb], [b.ptr.value for b in gt],
                       [b.ptr.value for b in gs])
    get = lambda bufs: [chunk_get(ctx, b.ptr.value, b.n) for b in bufs]  # noqa: E731
    all_phi, all_faces, all_gt, all_gs = get(phik), get(faces), get(gt), get(gs)
    for f in (2, 0, 1):
        sv.forward(f)
        phi = [chunk_get(ctx, p, n) for p, n in zip(sv.phi_ptrs(), sv.nphi)]
        fv = [chunk_get(ctx, p, n) for _, _, p, n in sv.exit_faces()]
        fbar = {(sd, nm): b.ptr.value for (sd, nm, _), b in zip(sv.exits, fb[f * ne:(f + 1) * ne])}
        sv.backward([b.ptr.value for b in wb[f * npr:(f + 1) * npr]], fbar, field=f)
        g1 = [chunk_get(ctx, p, n) for p, n in zip(sv.param_ptrs("grad_sigt"), sv.nparam)]
        g2 = [chunk_get(ctx, p, n) for p, n in zip(sv.param_ptrs("grad_s"), sv.nparam)]
        for x, y in zip(phi + fv + g1 + g2, all_phi[f * npr:(f + 1) * npr] + all_faces[f * ne:(f + 1) * ne] +
                        all_gt[f * npr:(f + 1) * npr] + all_gs[f * npr:(f + 1) * npr]):
            assert np.array_equal(x, y), "field %d differs from its own run" % f
    assert not np.array_equal(all_gt[0], all_gt[npr])  # the fields really differ
    sv.set_params([b.ptr.value for b in keep[0:npr]], [b.ptr.value for b in keep[npr:2 * npr]], field=1)
    with pytest.raises(ka.KripkeError):
        sv.backward(field=1)
    with pytest.raises(ka.KripkeError):
        sv.forward_fields(1, F)  # range beyond the fields
    for b in keep + phik + faces + wb + fb + gt + gs:
        b.free()
    sv.close()


@pytest.mark.parametrize("layout", ["DZG", "GZD", "ZDG", "ZGD"])
@pytest.mark.parametrize("kernel,zones,zset,groups,quad", [((2, 3), (16, 8, 32), (2, 1, 1), 32, (4, 4)),
                                                           ((2, 3), (32, 16, 12), (2, 2, 2), 64, (4, 2)),
                                                           ((3, 4), (32, 16, 32), (1, 1, 1), 64, (4, 8)),
                                                           ((3, 4), (64, 16, 64), (2, 1, 2), 32, (4, 8))])
def test_sweep_transpose_wave_kernel(ctx, layout, zones, zset, groups, quad, kernel):
    """The seeded transpose identity with the wave-uniform plane kernel, and
    with the systolic kernel keeping the faces between k-patches in LDS,
    installed (kr_problem_set_tuning): forward and adjoint sweeps both run it."""
    hp = product_for(ctx, **kwargs_for(layout, zset=zset, groups=groups, zones=zones, quad=quad, legendre=1))
    k = hp.kernels()
    assert k.set_tuning(v_max=kernel[0]) == kernel[1]
    rng = np.random.default_rng(48)
    exits = [(s, PLANES[a]) for s in range(hp.num_sdoms) for a, e in enumerate(k.sdom_exits(s)) if e]
    x = _rand_fields(hp, "rhs", hp.num_sdoms, rng)
    k.sweep_solver(list(range(hp.num_sdoms)))
    Hx = _get(hp, "psi", hp.num_sdoms)
    Bx = [hp.get(name, s) for s, name in exits]
    y = _rand_fields(hp, "psi", hp.num_sdoms, rng)
    fb = []
    for s, name in exits:
        v = rng.standard_normal(hp.get(name, s).size)
        hp.set(name, s, v)
        fb.append(v)
    k.sweep_solver_adjoint(seeded=True)
    Aty = _get(hp, "rhs", hp.num_sdoms)
    _check_dot((y + fb, Hx + Bx), (Aty, x))


@pytest.mark.parametrize("layout", ["DGZ", "GDZ"])
@pytest.mark.parametrize("zones,zset", [((32, 8, 12), (1, 1, 1)), ((16, 12, 20), (2, 2, 2))])
def test_sweep_transpose_four_zones_per_lane(ctx, layout, zones, zset):
    """The seeded transpose identity with the row-scan kernel at four zones
    per lane installed (kr_problem_set_tuning v_max = 4): forward and adjoint
    sweeps both run it."""
    hp = product_for(ctx, **kwargs_for(layout, zset=zset, groups=16, zones=zones, quad=(4, 4), legendre=1))
    k = hp.kernels()
    assert k.set_tuning(v_max=4) == 1
    rng = np.random.default_rng(47)
    exits = [(s, PLANES[a]) for s in range(hp.num_sdoms) for a, e in enumerate(k.sdom_exits(s)) if e]
    x = _rand_fields(hp, "rhs", hp.num_sdoms, rng)
    k.sweep_solver(list(range(hp.num_sdoms)))
    Hx = _get(hp, "psi", hp.num_sdoms)
    Bx = [hp.get(name, s) for s, name in exits]
    y = _rand_fields(hp, "psi", hp.num_sdoms, rng)
    fb = []
    for s, name in exits:
        v = rng.standard_normal(hp.get(name, s).size)
        hp.set(name, s, v)
        fb.append(v)
    k.sweep_solver_adjoint(seeded=True)
    Aty = _get(hp, "rhs", hp.num_sdoms)
    _check_dot((y + fb, Hx + Bx), (Aty, x))


@pytest.mark.parametrize("layout", ["DZG", "GZD", "ZDG", "ZGD"])
@pytest.mark.parametrize("kernel,zones,zset,groups,quad", [((0, 0), (16, 8, 12), (2, 2, 1), 32, (4, 4)),
                                                           ((3, 4), (32, 16, 32), (1, 1, 1), 64, (4, 8))])
def test_sweep_transpose_varying_widths(ctx, layout, zones, zset, groups, quad, kernel):
    """The seeded transpose identity on a mesh whose zone widths vary along
    every axis (the systolic kernel's instance for whole table rows), plain
    and with the faces between k-patches in LDS."""
    hp = product_for(ctx, **kwargs_for(layout, zset=zset, groups=groups, zones=zones, quad=quad, legendre=1))
    k = hp.kernels()
    assert k.set_tuning(v_max=kernel[0]) == kernel[1]
    rng = np.random.default_rng(49)
    for ax, name in enumerate(("dx", "dy", "dz")):
        for c in range(zset[ax]):
            w = hp.get(name, c)
            hp.set(name, c, w * (0.6 + 0.8 * rng.random(w.size)))
    exits = [(s, PLANES[a]) for s in range(hp.num_sdoms) for a, e in enumerate(k.sdom_exits(s)) if e]
    x = _rand_fields(hp, "rhs", hp.num_sdoms, rng)
    k.sweep_solver(list(range(hp.num_sdoms)))
    Hx = _get(hp, "psi", hp.num_sdoms)
    Bx = [hp.get(name, s) for s, name in exits]
    y = _rand_fields(hp, "psi", hp.num_sdoms, rng)
    fb = []
    for s, name in exits:
        v = rng.standard_normal(hp.get(name, s).size)
        hp.set(name, s, v)
        fb.append(v)
    k.sweep_solver_adjoint(seeded=True)
    Aty = _get(hp, "rhs", hp.num_sdoms)
    _check_dot((y + fb, Hx + Bx), (Aty, x))
