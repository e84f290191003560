"""The differentiable workload of the gradient checks: 32^3 zones, 32 groups,
96 directions (Gauss-Legendre 8 polar x 12 azimuthal, 8 direction sets), P2
(9 moments), DGZ, one zone set, 10 source iterations from psi = 0
(SteadyStateSolver.cpp:44-104), phi checkpoints every 2.

Loss J = sum over a seeded 4^3-zone detector box, every moment and group, of
(phi_K - phi_target)^2, where phi_target is phi_K of the same problem with
sigma_t(g,z) scaled by a seeded factor in [0.8, 1.2] per zone and group.

Parameters are sigma_t(g,z) (the sigt_zonal field) and the per-zone
isotropic scattering s(g,z); s is set to the mixed-element average of the
generated sigma_s, so the forward model is the reference's own iteration
with its generated scattering.  The 64 checked parameters alternate between
sigma_t and s (parameters() says where they lie).

Shared by the GPU gradient test and by tests/golden/make_c3_fd.py, which
differences the CPU oracle's forward run at the same points.  sigma_t and s
arrays are in DGZ's sigt_zonal layout: entry = g*Z + z, z = (k*ny + j)*nx + i.
"""
import numpy as np

KW = dict(zones=(32, 32, 32), groups=32, quad=(8, 12), legendre=2, layout="DGZ", dset=8, gset=1, zset=(1, 1, 1),
          geometry=1, seed=20261202)
ITERATIONS, CHECKPOINT = 10, 2
MOMENTS = 9
NUM_PARAMS = 64
H_REL = 1e-6
SEED = 20261204


def shape():
    nx, ny, nz = KW["zones"]
    return nx, ny, nz, KW["groups"]


def detector_corner():
    rng = np.random.default_rng(SEED)
    nx = KW["zones"][0]
    return tuple(int(c) for c in rng.integers(0, nx - 4, size=3))


def zone_mask():
    """Boolean mask over zones z (= (k*ny + j)*nx + i) of the detector box."""
    nx, ny, nz, _ = shape()
    ci, cj, ck = detector_corner()
    m = np.zeros((nz, ny, nx), dtype=bool)
    m[ck:ck + 4, cj:cj + 4, ci:ci + 4] = True
    return m.ravel()


def phi_mask():
    """Mask over phi's DGZ layout [nm][g][z]."""
    nx, ny, nz, G = shape()
    return np.broadcast_to(zone_mask()[None, None, :], (MOMENTS, G, nx * ny * nz)).ravel()


def target_scale():
    """Per (g, z) factor of sigma_t in the target problem."""
    nx, ny, nz, G = shape()
    rng = np.random.default_rng(SEED + 1)
    return 1.0 + 0.2 * (2.0 * rng.random(G * nx * ny * nz) - 1.0)


def zonal_scattering(z2m, z2n, mat, frac, sigs_diag):
    """s(g, z) = sum over the zone's mixed elements of fraction * sigma_s(mat, g)
    in mixed-element order (the order of the structure-aware scattering);
    sigs_diag[mat][g] is the n = 0 diagonal of data/sigs."""
    G = sigs_diag.shape[1]
    s = np.zeros((G, z2m.size))
    for z in range(z2m.size):
        acc = np.zeros(G)
        for q in range(z2n[z]):
            m = z2m[z] + q
            acc = acc + sigs_diag[mat[m]] * frac[m]
        s[:, z] = acc
    return s.ravel()


def parameters():
    """[(which, entry)] with which 0 = sigma_t, 1 = s.  Half lie in the
    detector box, 12 in pairs within two zones around it, 4 anywhere in the
    mesh (the box is optically thick at these cross sections, so far from it
    the gradient sinks toward the differences' rounding noise)."""
    nx, ny, nz, G = shape()
    rng = np.random.default_rng(SEED + 2)
    c = np.array(detector_corner())
    out = []
    for p in range(NUM_PARAMS):
        which, slot = p % 2, p // 2
        if slot < 16:
            lo, hi = c, c + 4
        elif slot < 28:
            lo, hi = np.clip(c - 2, 0, None), np.minimum(c + 6, [nx, ny, nz])
        else:
            lo, hi = np.zeros(3, dtype=int), np.array([nx, ny, nz])
        i, j, k = (int(rng.integers(lo[a], hi[a])) for a in range(3))
        g = int(rng.integers(0, G))
        out.append((which, g * nx * ny * nz + (k * ny + j) * nx + i))
    return out


def loss(phi, target):
    d = (phi - target)[phi_mask()]
    return float(np.dot(d, d))


def noise(J0, h):
    """Rounding noise of a central difference: J carries a few ulps of
    rounding in each evaluation, divided by 2h (32 ulps of J over h)."""
    return 32 * np.finfo(np.float64).eps * J0 / h


def tolerance(fd, J0, h):
    """Central differences: truncation O(h^2) inside 1e-6 relative, plus the
    rounding noise."""
    return 1e-6 * abs(fd) + noise(J0, h)
