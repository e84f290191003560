// Diamond-difference sweep for the zone-fastest nestings (DGZ, GDZ), where a
// zone row (fixed direction, group, j, k; i = 0..ni-1) is contiguous in HBM.
//
// A row group of RG lanes (RG = power of two >= ni) owns one k-row of the
// subdomain and walks its j-rows in sweep order, one row per step; row group
// kk runs kk steps behind row group 0 (wavefront over (j,k)), and a unit
// strings several (direction, group) pairs back to back so the fill is paid
// once.  Within a row the i-recurrence of the incoming x-face,
//     f_in(i+1) = 2 psi_i - f_in(i),   psi_i = (q_i + a_i f_in(i)) / den_i,
// is the affine map f -> alpha_i f + beta_i with alpha_i = 2 a_i/den_i - 1,
// beta_i = 2 q_i/den_i, solved for all lanes at once by an inclusive scan of
// map compositions (Hillis-Steele, log2 RG shuffle steps).  Every HBM access
// of rhs, sigma_t and psi is then one contiguous row per row group.
//
// The j-face stays in a register (same lane, next row); the k-face goes to
// row group kk+1 through LDS (double-buffered, one barrier per step).  The
// boundary faces (i-face per row, j-face row at j=0, k-face row at k=0) and
// rhs / sigma_t rows are prefetched RING steps ahead into registers.
//
// The scan re-associates the x-face recurrence; results differ from the
// sequential order by a few ulps per row (normwise ~1e-15).
#include "kr_launch.h"

namespace kr {

namespace {

constexpr int RRING = 8;

struct RowCursor {
  int32_t jj, r;  // row within the item (sweep order), pair within the unit
};

__device__ inline void row_advance(RowCursor &c, int nj) {
  if (++c.jj == nj) {
    c.jj = 0;
    ++c.r;
  }
}

struct RowItem {
  int32_t d, g;
  int64_t pbase;  // d*fa + g*fg
};

__device__ inline RowItem row_item(const DevSdom &S, int inner_is_g, int pidx) {
  RowItem it;
  if (inner_is_g) {
    it.d = pidx / S.ng;
    it.g = pidx - it.d * S.ng;
  } else {
    it.g = pidx / S.nd;
    it.d = pidx - it.g * S.nd;
  }
  it.pbase = it.d * S.fa + it.g * S.fg;
  return it;
}

// Inclusive scan of affine maps f -> A f + B over the RG lanes of a row
// group; lane l ends with the composition of maps 0..l (map 0 applied first).
template <int RG>
__device__ inline void affine_scan(double &A, double &B, int l) {
#pragma unroll
  for (int off = 1; off < RG; off <<= 1) {
    const double Ap = __shfl_up(A, off, RG);
    const double Bp = __shfl_up(B, off, RG);
    if (l >= off) {
      B = __fma_rn(A, Bp, B);
      A = A * Ap;
    }
  }
}

}  // namespace

template <int RG>
__global__ __launch_bounds__(1024) void k_sweep_rows(const DevSdom *__restrict__ sd,
                                                     const SweepPlanes *__restrict__ planes,
                                                     const SweepUnit *__restrict__ units, SweepShape sh) {
  extern __shared__ double lds[];
  const int nt = blockDim.x;
  const SweepUnit U = units[blockIdx.x];
  const DevSdom &S = sd[U.sdom];
  const SweepPlanes &PL = planes[U.sdom];
  const int tid = threadIdx.x;
  const int l = tid % RG;   // position along the row in sweep order
  const int kk = tid / RG;  // k-row (sweep order) owned by this row group
  const int ni = S.ni, nj = S.nj, nk = S.nk, nd = S.nd;
  const bool live = l < ni;
  const int i = live ? (S.id > 0 ? l : ni - 1 - l) : 0;
  const int k = S.kd > 0 ? kk : nk - 1 - kk;
  const int R = U.pg_count;

  // LDS: k-face exchange [2][nt], then the 2cos/delta tables A[nd][ni] B[nd][nj] C[nd][nk]
  double *fzb = lds;
  double *cA = lds + 2 * nt;
  double *cB = cA + nd * ni;
  double *cC = cB + nd * nj;
  {
    const gmem<double> *src = gp(S.coef[0]);  // A, B, C are contiguous per subdomain
    for (int e = tid; e < nd * (ni + nj + nk); e += nt) cA[e] = src[e];
  }
  __syncthreads();

  const gmem<double> *rhs = gp(S.rhs);
  const gmem<double> *sig = gp(S.sig);
  gmem<double> *psi = gp(S.psi);
  const gmem<double> *in0 = gp(PL.in[0]), *in1 = gp(PL.in[1]), *in2 = gp(PL.in[2]);
  gmem<double> *out0 = gp(PL.out[0]), *out1 = gp(PL.out[1]), *out2 = gp(PL.out[2]);
  const int64_t zrow_k = (int64_t)k * nj;  // z = (k*nj + j)*ni + i

  double rq[RRING], rs[RRING], rz[RRING];
  // Faces needed once per item, loaded when the loader enters the item (RRING
  // steps early; nj >= RRING so at most one is pending): the x-faces of all
  // nj rows (lane l holds row j = l of i_plane) and the j = 0 face row.
  double xnext = 0.0, ynext = 0.0;

  auto load_slot = [&](int r, const RowCursor &c, const RowItem &it) {
    const int j = S.jd > 0 ? c.jj : nj - 1 - c.jj;
    const int64_t z = (zrow_k + j) * ni + i;
    rq[r] = live ? rhs[it.pbase + z * S.fz] : 0.0;
    rs[r] = live ? sig[it.g * S.sg + z * S.sz] : 1.0;
    if (kk == 0) rz[r] = (in2 && live) ? in2[it.d * S.pa[2] + it.g * S.pg[2] + ((int64_t)j * ni + i) * S.pf[2]] : 0.0;
    if (c.jj == 0) {
      xnext = (in0 && l < nj) ? in0[it.d * S.pa[0] + it.g * S.pg[0] + (zrow_k + l) * S.pf[0]] : 0.0;
      ynext = (in1 && live) ? in1[it.d * S.pa[1] + it.g * S.pg[1] + ((int64_t)k * ni + i) * S.pf[1]] : 0.0;
    }
  };

  RowCursor lc{-kk, 0};
  RowItem li = row_item(S, sh.inner_is_g, U.pg_begin);
#pragma unroll
  for (int r = 0; r < RRING; ++r) {
    if (lc.jj >= 0 && lc.r < R) {
      if (lc.jj == 0) li = row_item(S, sh.inner_is_g, U.pg_begin + lc.r);
      load_slot(r, lc, li);
    }
    row_advance(lc, nj);
  }

  RowCursor cc{-kk, 0};
  RowItem ci = row_item(S, sh.inner_is_g, U.pg_begin);
  double fy = 0.0, a = 0.0, ck = 0.0, xcur = 0.0;
  const int64_t T = (int64_t)R * nj + (sh.PK - 1);

  for (int64_t t0 = 0; t0 < T; t0 += RRING) {
#pragma unroll
    for (int r = 0; r < RRING; ++r) {
      if (t0 + r < T) {
        const int cur = r & 1, prev = cur ^ 1;
        if (cc.jj >= 0 && cc.r < R) {
          if (cc.jj == 0) {
            ci = row_item(S, sh.inner_is_g, U.pg_begin + cc.r);
            a = cA[ci.d * ni + i];
            ck = cC[ci.d * nk + k];
            fy = ynext;
            xcur = xnext;
          }
          const int j = S.jd > 0 ? cc.jj : nj - 1 - cc.jj;
          const double bj = cB[ci.d * nj + j];
          const double fz = (kk == 0) ? rz[r] : fzb[prev * nt + tid - RG];
          const double den = ((a + bj) + ck) + rs[r];
          const double inv = 1.0 / den;
          const double q = __fma_rn(fz, ck, __fma_rn(fy, bj, rq[r]));
          double A = live ? __fma_rn(2.0 * a, inv, -1.0) : 1.0;
          double B = live ? 2.0 * q * inv : 0.0;
          affine_scan<RG>(A, B, l);
          const double f0 = __shfl(xcur, j, RG);  // x-face entering row j
          const double fout = __fma_rn(A, f0, B);
          double fin = __shfl_up(fout, 1, RG);
          if (l == 0) fin = f0;
          const double p = __fma_rn(a, fin, q) * inv;
          const double p2 = 2.0 * p;
          const double fyo = p2 - fy, fzo = p2 - fz;
          fy = fyo;
          if (live) {
            const int64_t z = (zrow_k + j) * ni + i;
            psi[ci.pbase + z * S.fz] = p;
            if (l == ni - 1) out0[ci.d * S.pa[0] + ci.g * S.pg[0] + (zrow_k + j) * S.pf[0]] = p2 - fin;
            if (cc.jj == nj - 1) out1[ci.d * S.pa[1] + ci.g * S.pg[1] + ((int64_t)k * ni + i) * S.pf[1]] = fyo;
            if (kk == sh.PK - 1) out2[ci.d * S.pa[2] + ci.g * S.pg[2] + ((int64_t)j * ni + i) * S.pf[2]] = fzo;
          }
          fzb[cur * nt + tid] = fzo;
        }
        row_advance(cc, nj);
        if (lc.jj >= 0 && lc.r < R) {
          if (lc.jj == 0) li = row_item(S, sh.inner_is_g, U.pg_begin + lc.r);
          load_slot(r, lc, li);
        }
        row_advance(lc, nj);
        __syncthreads();
      }
    }
  }
}

int launch_sweep_rows(hipStream_t st, const DevSdom *sd, const SweepPlanes *planes, const SweepUnit *units,
                      int32_t num_units, const SweepShape &shape, int32_t nd, int32_t ni, int32_t nj, int32_t nk) {
  if (num_units <= 0) return KR_OK;
  const int nt = shape.RG * shape.PK;
  const size_t lds = (2 * (size_t)nt + (size_t)nd * (ni + nj + nk)) * sizeof(double);
  switch (shape.RG) {
#define X(rg)                                                                                                  \
  case rg:                                                                                                     \
    hipLaunchKernelGGL(k_sweep_rows<rg>, dim3(num_units), dim3(nt), lds, st, sd, planes, units, shape);        \
    break;
    X(8) X(16) X(32) X(64)
#undef X
    default: set_error("row sweep: unsupported row-group width"); return KR_ERR_UNSUPPORTED;
  }
  KR_HIP_CHECK(hipGetLastError());
  return KR_OK;
}

}  // namespace kr
