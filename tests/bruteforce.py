"""Independent numpy formulations used to pin the oracle on tiny problems.

None of these follows the oracle's loop structure: LTimes/LPlusTimes are
einsum contractions, the sweep is the diamond-difference discretisation
written as one global linear system (zone balance + face closure + vacuum
inflow) and solved densely, and sums use math.fsum.
"""
import math

import numpy as np

from helpers import LAYOUT_AXES, canonical


def ltimes_numpy(o, layout, M):
    """phi(nm,g,z) = sum over subdomains sharing (gset, zset), directions of
    ell(d,nm) psi(d,g,z)."""
    out = {}
    for s in range(o.num_sdoms):
        S = o.sdom(s)
        nd, ng, nz = S["nd"], S["ng"], S["ni"] * S["nj"] * S["nk"]
        psi = canonical(o.get("psi", s), layout, nd, ng, nz)
        ell = o.get("ell", S["dset"]).reshape(nd, M)
        contrib = np.einsum("dm,dgz->mgz", ell, psi)
        out[S["pr"]] = out.get(S["pr"], 0.0) + contrib
    return out


def lplustimes_numpy(o, layout, M):
    out = {}
    for s in range(o.num_sdoms):
        S = o.sdom(s)
        nd, ng, nz = S["nd"], S["ng"], S["ni"] * S["nj"] * S["nk"]
        po = canonical(o.get("phi_out", S["pr"]), layout, M, ng, nz)
        elp = o.get("ell_plus", S["dset"]).reshape(nd, M)
        out[s] = np.einsum("dm,mgz->dgz", elp, po)
    return out


def dd_sweep_linear_system(q, sigt, a_i, b_j, c_k, signs):
    """Solve the diamond-difference equations of one (direction, group) on an
    ni x nj x nk block with vacuum inflow, as one linear system.

    Unknowns: psi per zone and every face value.  Equations per zone:
      (a+b+c+sigt) psi - a Fx_in - b Fy_in - c Fz_in = q
      Fx_out + Fx_in = 2 psi   (and the same for y, z)
    plus Fx_in = 0 on the upwind boundary faces.
    Arrays q, sigt are [k][j][i]; a_i[i] = 2|mu|/dx_i etc.; signs = (id, jd, kd).
    Returns psi[k][j][i].
    """
    nk, nj, ni = q.shape
    nz = ni * nj * nk
    nfx = (ni + 1) * nj * nk
    nfy = ni * (nj + 1) * nk
    nfz = ni * nj * (nk + 1)
    n = nz + nfx + nfy + nfz

    def zid(i, j, k):
        return (k * nj + j) * ni + i

    def fx(f, j, k):  # x-face f in 0..ni
        return nz + (k * nj + j) * (ni + 1) + f

    def fy(i, f, k):
        return nz + nfx + (k * (nj + 1) + f) * ni + i

    def fz(i, j, f):
        return nz + nfx + nfy + (f * nj + j) * ni + i

    A = np.zeros((n, n))
    rhs = np.zeros(n)
    row = 0
    idn, jdn, kdn = signs
    for k in range(nk):
        for j in range(nj):
            for i in range(ni):
                z = zid(i, j, k)
                xin, xout = (fx(i, j, k), fx(i + 1, j, k)) if idn > 0 else (fx(i + 1, j, k), fx(i, j, k))
                yin, yout = (fy(i, j, k), fy(i, j + 1, k)) if jdn > 0 else (fy(i, j + 1, k), fy(i, j, k))
                zin, zout = (fz(i, j, k), fz(i, j, k + 1)) if kdn > 0 else (fz(i, j, k + 1), fz(i, j, k))
                a, b, c = a_i[i], b_j[j], c_k[k]
                A[row, z] = a + b + c + sigt[k, j, i]
                A[row, xin] = -a
                A[row, yin] = -b
                A[row, zin] = -c
                rhs[row] = q[k, j, i]
                row += 1
                for fin, fout in ((xin, xout), (yin, yout), (zin, zout)):
                    A[row, fout] = 1.0
                    A[row, fin] = 1.0
                    A[row, z] = -2.0
                    row += 1
    # vacuum inflow faces
    for k in range(nk):
        for j in range(nj):
            A[row, fx(0 if idn > 0 else ni, j, k)] = 1.0
            row += 1
    for k in range(nk):
        for i in range(ni):
            A[row, fy(i, 0 if jdn > 0 else nj, k)] = 1.0
            row += 1
    for j in range(nj):
        for i in range(ni):
            A[row, fz(i, j, 0 if kdn > 0 else nk)] = 1.0
            row += 1
    assert row == n
    sol = np.linalg.solve(A, rhs)
    return sol[:nz].reshape(nk, nj, ni)


def global_sweep_numpy(o, layout, nx, ny, nz_):
    """Sweep every (direction, group) of the whole domain (all zone sets
    merged) from the oracle's rhs, sigt_zonal, quadrature and mesh; returns
    psi per subdomain in canonical [d][g][z] order."""
    S0 = o.sdom(0)
    nd, ng = S0["nd"], S0["ng"]
    # global rhs / sigt arrays indexed [dset][d][g][k][j][i] and [gset][g][k][j][i]
    sd = [o.sdom(s) for s in range(o.num_sdoms)]
    nsets_d = max(S["dset"] for S in sd) + 1
    nsets_g = max(S["gset"] for S in sd) + 1
    Q = np.zeros((nsets_d, nsets_g, nd, ng, nz_, ny, nx))
    SIG = np.zeros((nsets_g, ng, nz_, ny, nx))
    for s, S in enumerate(sd):
        ni, nj, nk = S["ni"], S["nj"], S["nk"]
        r = canonical(o.get("rhs", s), layout, nd, ng, ni * nj * nk).reshape(nd, ng, nk, nj, ni)
        Q[S["dset"], S["gset"], :, :, S["klo"]:S["klo"] + nk, S["jlo"]:S["jlo"] + nj, S["ilo"]:S["ilo"] + ni] = r
        st = o.get("sigt_zonal", S["pr"])
        # sigt_zonal is [g][z] when Group precedes Zone in the nesting, else [z][g]
        gz_first = layout in ("DGZ", "GDZ", "GZD")
        stc = st.reshape(ng, -1) if gz_first else st.reshape(-1, ng).T
        SIG[S["gset"], :, S["klo"]:S["klo"] + nk, S["jlo"]:S["jlo"] + nj, S["ilo"]:S["ilo"] + ni] = \
            stc.reshape(ng, nk, nj, ni)
    dx, dy, dz = o.get("dx"), o.get("dy"), o.get("dz")
    out = {}
    for q in range(nsets_d):
        xc, yc, zc = o.get("xcos", q), o.get("ycos", q), o.get("zcos", q)
        sig_id, sig_jd, sig_kd = o.get("id", q), o.get("jd", q), o.get("kd", q)
        for gs in range(nsets_g):
            psi = np.zeros((nd, ng, nz_, ny, nx))
            for d in range(nd):
                a = 2.0 * xc[d] / dx
                b = 2.0 * yc[d] / dy
                c = 2.0 * zc[d] / dz
                signs = (int(sig_id[0]), int(sig_jd[0]), int(sig_kd[0]))
                for g in range(ng):
                    psi[d, g] = dd_sweep_linear_system(Q[q, gs, d, g], SIG[gs, g], a, b, c, signs)
            out[(q, gs)] = psi
    res = {}
    for s, S in enumerate(sd):
        ni, nj, nk = S["ni"], S["nj"], S["nk"]
        p = out[(S["dset"], S["gset"])][:, :, S["klo"]:S["klo"] + nk, S["jlo"]:S["jlo"] + nj, S["ilo"]:S["ilo"] + ni]
        res[s] = p.reshape(nd, ng, nk * nj * ni)
    return res


def blockwise_sweep_numpy(o, layout):
    """Block Jacobi on one rank: every subdomain solved alone with vacuum
    inflow on all of its faces, from the oracle's rhs, sigt_zonal, quadrature
    and mesh; psi per subdomain in canonical [d][g][z] order."""
    dx, dy, dz = o.get("dx"), o.get("dy"), o.get("dz")
    gz_first = layout in ("DGZ", "GDZ", "GZD")
    res = {}
    for s in range(o.num_sdoms):
        S = o.sdom(s)
        nd, ng, ni, nj, nk = S["nd"], S["ng"], S["ni"], S["nj"], S["nk"]
        q = canonical(o.get("rhs", s), layout, nd, ng, ni * nj * nk).reshape(nd, ng, nk, nj, ni)
        st = o.get("sigt_zonal", S["pr"])
        sig = (st.reshape(ng, -1) if gz_first else st.reshape(-1, ng).T).reshape(ng, nk, nj, ni)
        xc, yc, zc = o.get("xcos", S["dset"]), o.get("ycos", S["dset"]), o.get("zcos", S["dset"])
        signs = tuple(int(o.get(n, S["dset"])[0]) for n in ("id", "jd", "kd"))
        ddx = dx[S["ilo"]:S["ilo"] + ni]
        ddy = dy[S["jlo"]:S["jlo"] + nj]
        ddz = dz[S["klo"]:S["klo"] + nk]
        psi = np.zeros((nd, ng, nk, nj, ni))
        for d in range(nd):
            for g in range(ng):
                psi[d, g] = dd_sweep_linear_system(q[d, g], sig[g], 2.0 * xc[d] / ddx, 2.0 * yc[d] / ddy,
                                                   2.0 * zc[d] / ddz, signs)
        res[s] = psi.reshape(nd, ng, nk * nj * ni)
    return res


def population_fsum(o, layout):
    terms = []
    for s in range(o.num_sdoms):
        S = o.sdom(s)
        nd, ng, nz = S["nd"], S["ng"], S["ni"] * S["nj"] * S["nk"]
        psi = canonical(o.get("psi", s), layout, nd, ng, nz)
        w = o.get("w", S["dset"])
        vol = o.get("volume", S["r"])
        t = (w[:, None, None] * psi) * vol[None, None, :]
        terms.append(t.ravel())
    return math.fsum(np.concatenate(terms))


def mix_maps(o, r):
    """Mixed-element maps of zone-set chunk r as integer / float arrays."""
    return (o.get("zone_to_mixelem", r).astype(int), o.get("zone_to_num_mixelem", r).astype(int),
            o.get("mixelem_to_material", r).astype(int), o.get("mixelem_to_fraction", r))


def numpy_iteration(o, layout, M, m2l, strength=1.0):
    """One source iteration from the oracle's current psi, in numpy."""
    S0 = o.sdom(0)
    ng, nz, nd = S0["ng"], S0["ni"] * S0["nj"] * S0["nk"], S0["nd"]
    phi = ltimes_numpy(o, layout, M)
    sd = [o.sdom(s) for s in range(o.num_sdoms)]
    sigs = o.get("sigs").reshape(3, -1, o.inp.num_groups, o.inp.num_groups)
    phi_out = {}
    for S in sd:
        if S["dset"] != 0:
            continue
        z2m, z2n, mat, frac = mix_maps(o, S["r"])
        sz = np.zeros((nz, sigs.shape[1], ng, ng))
        G0 = S["glo"]
        for z in range(nz):
            for k in range(z2n[z]):
                sz[z] += frac[z2m[z] + k] * sigs[mat[z2m[z] + k], :, G0:G0 + ng, G0:G0 + ng]
        po_ = np.einsum("zmgh,mhz->mgz", sz[:, m2l], phi[S["pr"]])
        for k in range(len(mat)):
            if mat[k] == 0:
                po_[0, :, int(o.get("mixelem_to_zone", S["r"])[k])] += strength * frac[k]
        phi_out[S["pr"]] = po_
    for c, v in phi_out.items():
        o.set("phi_out", c, np.transpose(v, [("agz").index(ch) for ch in LAYOUT_AXES[layout]]).ravel())
    rhs = lplustimes_numpy(o, layout, M)
    for s, v in rhs.items():
        o.set("rhs", s, np.transpose(v.reshape(nd, ng, nz),
                                     [("agz").index(ch) for ch in LAYOUT_AXES[layout]]).ravel())
    psi = global_sweep_numpy(o, layout, o.inp.nx, o.inp.ny, o.inp.nz)
    return phi, psi
