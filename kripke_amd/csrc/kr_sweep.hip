// Diamond-difference sweep (src/Kripke/Kernel/SweepSubdomain.cpp:26-110)
// as a systolic pencil wavefront on gfx950.
//
// A workgroup owns a patch of (j,k) zone columns of one subdomain, times Q
// adjacent (direction, group) pairs.  Each thread walks its i-pencil one
// zone per step; the thread at (jj,kk) runs (jj+kk) steps behind the patch
// corner, so zone (i,j,k) is computed at step i+j+k after its upwind
// neighbours (i-1,j,k), (i,j-1,k), (i,j,k-1).  The i-face stays in a
// register, the j- and k-faces pass to the neighbouring thread through LDS
// (double-buffered, one barrier per step).  A unit strings several pair
// groups and patches back to back through the same pipeline, so the
// wavefront fill is paid once per unit.  Faces that leave a patch go to the
// subdomain's own plane chunk (the reference's running-plane storage);
// faces that leave the subdomain go to the downwind subdomain's plane chunk
// (pipeline mode) or stay in the own chunk (Kernel::sweepSubdomain mode).
//
// Arithmetic per zone (bit-identical across nestings and schedules):
//   num = fma(fz, c_k, fma(fy, b_j, fma(fx, a_i, rhs)))
//   psi = num / (((a_i + b_j) + c_k) + sigt)          (IEEE division)
//   f_out = 2*psi - f_in                              (exact scaling)
// with a_i = 2*xcos(d)/dx(i) etc. precomputed on the host exactly as the
// reference computes them per zone.
#include "kr_launch.h"

#include <algorithm>
#include <cstdlib>

namespace kr {

constexpr int RING = 8;  // prefetch depth (steps) of rhs / sigt / a_i

struct Cursor {
  int32_t ii, pj, pk, r;
};

__device__ inline void advance(Cursor &c, int ni, int NPJ, int NPK) {
  if (++c.ii == ni) {
    c.ii = 0;
    if (++c.pj == NPJ) {
      c.pj = 0;
      if (++c.pk == NPK) {
        c.pk = 0;
        ++c.r;
      }
    }
  }
}

struct ItemState {
  int32_t d, g, j, k;
  int64_t fbase, sbase;
};

__device__ inline ItemState item_state(const DevSdom &S, const SweepShape &sh, int pg, int q, int jj, int kk,
                                       const Cursor &c) {
  ItemState st;
  const int pidx = pg * sh.Q + q;
  if (sh.inner_is_g) {
    st.d = pidx / S.ng;
    st.g = pidx - st.d * S.ng;
  } else {
    st.g = pidx / S.nd;
    st.d = pidx - st.g * S.nd;
  }
  const int jl = c.pj * sh.PJ + jj, kl = c.pk * sh.PK + kk;
  st.j = S.jd > 0 ? jl : S.nj - 1 - jl;
  st.k = S.kd > 0 ? kl : S.nk - 1 - kl;
  const int64_t zrow = ((int64_t)st.k * S.nj + st.j) * S.ni;
  st.fbase = st.d * S.fa + st.g * S.fg + zrow * S.fz;
  st.sbase = st.g * S.sg + zrow * S.sz;
  return st;
}

__global__ __launch_bounds__(1024) void k_sweep(const DevSdom *__restrict__ sd, const SweepPlanes *__restrict__ planes,
                                                const SweepUnit *__restrict__ units, SweepShape sh) {
  extern __shared__ double lds[];
  const int nt = blockDim.x;
  double *fyb = lds;           // [2][nt]
  double *fzb = lds + 2 * nt;  // [2][nt]
  const SweepUnit U = units[blockIdx.x];
  const DevSdom &S = sd[U.sdom];
  const SweepPlanes &PL = planes[U.sdom];
  const int tid = threadIdx.x;
  const int Q = sh.Q, PJ = sh.PJ, PK = sh.PK, NPJ = sh.NPJ, NPK = sh.NPK;
  const int q = tid % Q, jj = (tid / Q) % PJ, kk = tid / (Q * PJ);
  const int s = jj + kk;
  const int ni = S.ni, R = U.pg_count;
  const int64_t NU = (int64_t)R * NPJ * NPK * ni;
  const int64_t T = NU + (PJ - 1) + (PK - 1);
  const gmem<double> *coefA = gp(S.coef[0]);
  const gmem<double> *coefB = gp(S.coef[1]);
  const gmem<double> *coefC = gp(S.coef[2]);
  const gmem<double> *rhs = gp(S.rhs);
  const gmem<double> *sig = gp(S.sig);
  gmem<double> *psi = gp(S.psi);
  const gmem<double> *in0 = gp(PL.in[0]), *in1 = gp(PL.in[1]), *in2 = gp(PL.in[2]);
  gmem<double> *out0 = gp(PL.out[0]), *out1 = gp(PL.out[1]), *out2 = gp(PL.out[2]);
  gmem<double> *own1 = gp(S.plane[1]), *own2 = gp(S.plane[2]);
  const int64_t fz = S.fz, sz = S.sz;

  double ring_q[RING], ring_s[RING], ring_a[RING];

  // Loader cursor starts at local time u = -s and runs RING steps ahead.
  Cursor lc{-s, 0, 0, 0};
  ItemState ls{};
  bool ls_valid = false;
#pragma unroll
  for (int r = 0; r < RING; ++r) {
    if (lc.ii >= 0 && lc.r < R) {
      if (!ls_valid || lc.ii == 0) {
        ls = item_state(S, sh, U.pg_begin + lc.r, q, jj, kk, lc);
        ls_valid = true;
      }
      const int i = S.id > 0 ? lc.ii : ni - 1 - lc.ii;
      ring_q[r] = rhs[ls.fbase + i * fz];
      ring_s[r] = sig[ls.sbase + i * sz];
      ring_a[r] = coefA[ls.d * ni + i];
    }
    advance(lc, ni, NPJ, NPK);
  }

  Cursor cc{-s, 0, 0, 0};
  ItemState cs{};
  double fx = 0.0, bj = 0.0, ck = 0.0;

  for (int64_t t0 = 0; t0 < T; t0 += RING) {
#pragma unroll
    for (int r = 0; r < RING; ++r) {
      if (t0 + r < T) {
        const int cur = r & 1, prev = cur ^ 1;
        if (cc.ii >= 0 && cc.r < R) {
          if (cc.ii == 0) {
            cs = item_state(S, sh, U.pg_begin + cc.r, q, jj, kk, cc);
            bj = coefB[cs.d * S.nj + cs.j];
            ck = coefC[cs.d * S.nk + cs.k];
            const gmem<double> *src = in0;
            fx = src ? src[cs.d * S.pa[0] + cs.g * S.pg[0] + ((int64_t)cs.k * S.nj + cs.j) * S.pf[0]] : 0.0;
          }
          const int i = S.id > 0 ? cc.ii : ni - 1 - cc.ii;
          double fy, fzv;
          if (jj == 0) {
            const gmem<double> *src = (cc.pj == 0) ? in1 : own1;
            fy = src ? src[cs.d * S.pa[1] + cs.g * S.pg[1] + ((int64_t)cs.k * ni + i) * S.pf[1]] : 0.0;
          } else {
            fy = fyb[prev * nt + tid - Q];
          }
          if (kk == 0) {
            const gmem<double> *src = (cc.pk == 0) ? in2 : own2;
            fzv = src ? src[cs.d * S.pa[2] + cs.g * S.pg[2] + ((int64_t)cs.j * ni + i) * S.pf[2]] : 0.0;
          } else {
            fzv = fzb[prev * nt + tid - Q * PJ];
          }
          const double a = ring_a[r];
          const double num = __fma_rn(fzv, ck, __fma_rn(fy, bj, __fma_rn(fx, a, ring_q[r])));
          const double den = ((a + bj) + ck) + ring_s[r];
          const double p = num / den;
          psi[cs.fbase + i * fz] = p;
          const double p2 = 2.0 * p;
          fx = p2 - fx;
          const double fyo = p2 - fy, fzo = p2 - fzv;
          if (cc.ii == ni - 1)
            out0[cs.d * S.pa[0] + cs.g * S.pg[0] + ((int64_t)cs.k * S.nj + cs.j) * S.pf[0]] = fx;
          if (jj == PJ - 1) {
            gmem<double> *dst = (cc.pj == NPJ - 1) ? out1 : own1;
            dst[cs.d * S.pa[1] + cs.g * S.pg[1] + ((int64_t)cs.k * ni + i) * S.pf[1]] = fyo;
          } else {
            fyb[cur * nt + tid] = fyo;
          }
          if (kk == PK - 1) {
            gmem<double> *dst = (cc.pk == NPK - 1) ? out2 : own2;
            dst[cs.d * S.pa[2] + cs.g * S.pg[2] + ((int64_t)cs.j * ni + i) * S.pf[2]] = fzo;
          } else {
            fzb[cur * nt + tid] = fzo;
          }
        }
        advance(cc, ni, NPJ, NPK);
        // refill this ring slot with the values for step t + RING
        if (lc.ii >= 0 && lc.r < R) {
          if (!ls_valid || lc.ii == 0) {
            ls = item_state(S, sh, U.pg_begin + lc.r, q, jj, kk, lc);
            ls_valid = true;
          }
          const int i = S.id > 0 ? lc.ii : ni - 1 - lc.ii;
          ring_q[r] = rhs[ls.fbase + i * fz];
          ring_s[r] = sig[ls.sbase + i * sz];
          ring_a[r] = coefA[ls.d * ni + i];
        }
        advance(lc, ni, NPJ, NPK);
        __syncthreads();
      }
    }
  }
}

// 2*cos(d)/delta per (direction, zone index), the exact expression of
// SweepSubdomain.cpp:79-81 (product by 2 is exact, IEEE division), for the
// subdomains [first, first+count).  Recomputed before every sweep so that
// the quadrature and mesh fields are read as they are at call time.
__global__ __launch_bounds__(256) void k_sweep_coef(const DevSdom *__restrict__ sd, int first) {
  const DevSdom &S = sd[first + blockIdx.y];
  const int n[3] = {S.ni, S.nj, S.nk};
  const int total = S.nd * (n[0] + n[1] + n[2]);
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < total; e += gridDim.x * blockDim.x) {
    const int d = e / (n[0] + n[1] + n[2]);
    int r = e - d * (n[0] + n[1] + n[2]);
    int ax = 0;
    while (r >= n[ax]) {
      r -= n[ax];
      ++ax;
    }
    gp(S.coef[ax])[d * n[ax] + r] = 2.0 * gp(S.cosines[ax])[d] / gp(S.delta[ax])[r];
  }
}

int launch_sweep_coef(hipStream_t st, const DevSdom *sd, int first, int count) {
  if (count <= 0) return KR_OK;
  hipLaunchKernelGGL(k_sweep_coef, dim3(4, count), dim3(256), 0, st, sd, first);
  KR_HIP_CHECK(hipGetLastError());
  return KR_OK;
}

static int largest_divisor_le(int n, int cap) {
  for (int d = cap < n ? cap : n; d >= 1; --d)
    if (n % d == 0) return d;
  return 1;
}

// Row-scan shape for zone-fastest nestings: V zones per lane (V > 1 needs
// 16-byte aligned rows), W >= ni/V lanes per row group, one row group per
// k-row.  Returns false when the subdomain does not fit one workgroup.
static bool row_shape(SweepShape &sh, int32_t nd, int32_t ng, int32_t ni, int32_t nj, int32_t nk, bool aligned16) {
  int want = ni % 2 == 0 ? 2 : 1;
  if (const char *e = getenv("KRIPKE_AMD_ROW_ZONES")) want = atoi(e);  // tuning override: 1 or 2
  const int64_t nz = (int64_t)ni * nj * nk;
  const int64_t maxface = std::max({(int64_t)nj * nk, (int64_t)ni * nk, (int64_t)ni * nj});
  // every stream is a buffer resource addressed with 32-bit byte offsets below 2 GiB
  if ((int64_t)nd * ng * nz * 8 >= (1LL << 31) || (int64_t)nd * ng * maxface * 8 >= (1LL << 31)) return false;
  for (int V = 2; V >= 1; V /= 2) {
    if (V > want || ni % V != 0 || (V > 1 && !aligned16)) continue;
    int W = 4;
    while (W * V < ni) W *= 2;
    const size_t lds = (2 * (size_t)W * nk * V + (size_t)nd * (ni + nj + nk + 1)) * sizeof(double);
    if (W <= 64 && W * nk <= 1024 && lds <= 65536) {
      sh.kind = 1;
      sh.RG = W;
      sh.V = V;
      sh.Q = 1;
      sh.PJ = nj;
      sh.PK = nk;
      sh.NPJ = 1;
      sh.NPK = 1;
      return true;
    }
  }
  return false;
}

SweepShape choose_sweep_shape(int nesting, int32_t nd, int32_t ng, int32_t ni, int32_t nj, int32_t nk, int ref_mode,
                              bool aligned16) {
  SweepShape sh{};
  sh.ref_mode = ref_mode;
  sh.nd = nd;
  sh.ni = ni;
  sh.nj = nj;
  sh.nk = nk;
  const bool zfast = (nesting == KR_DGZ || nesting == KR_GDZ);
  // inner pair index = the faster of (d, g) in psi's memory order
  sh.inner_is_g = (nesting == KR_DGZ || nesting == KR_DZG || nesting == KR_ZDG) ? 1 : 0;
  const int inner = sh.inner_is_g ? ng : nd;
  if (zfast && row_shape(sh, nd, ng, ni, nj, nk, aligned16)) return sh;
  if (zfast) {
    sh.Q = 1;
  } else {
    int qv = 1;
    while (qv < 16 && inner % (qv * 2) == 0) qv *= 2;
    sh.Q = qv;
  }
  const int cap = 1024 / sh.Q;
  // patch: as many (j,k) columns as fit, subject to the pipeline distances
  // (a patch's downstream neighbour must start after its face is written).
  int best_pj = 1, best_pk = 1;
  for (int pj = largest_divisor_le(nj, cap); pj >= 1; pj = largest_divisor_le(nj, pj - 1)) {
    const int pk = largest_divisor_le(nk, cap / pj);
    const int npj = nj / pj;
    const bool ok_j = (npj == 1) || (ni >= pj);
    const bool ok_k = (nk / pk == 1) || ((int64_t)npj * ni >= pk);
    if (ok_j && ok_k && pj * pk > best_pj * best_pk) {
      best_pj = pj;
      best_pk = pk;
    }
    if (pj == 1) break;
  }
  sh.PJ = best_pj;
  sh.PK = best_pk;
  sh.NPJ = nj / best_pj;
  sh.NPK = nk / best_pk;
  return sh;
}

int launch_sweep(hipStream_t st, const DevSdom *sd, const SweepPlanes *planes, const SweepUnit *units,
                 int32_t num_units, const SweepShape &shape, double *pop) {
  if (num_units <= 0) return KR_OK;
  if (shape.kind == 1)
    return launch_sweep_rows(st, sd, planes, units, num_units, shape, shape.nd, shape.ni, shape.nj, shape.nk, pop);
  if (pop) {
    set_error("sweep: the systolic kernel has no fused population");
    return KR_ERR_UNSUPPORTED;
  }
  const int nt = shape.Q * shape.PJ * shape.PK;
  const size_t lds = 4 * (size_t)nt * sizeof(double);
  hipLaunchKernelGGL(k_sweep, dim3(num_units), dim3(nt), lds, st, sd, planes, units, shape);
  KR_HIP_CHECK(hipGetLastError());
  return KR_OK;
}

}  // namespace kr
