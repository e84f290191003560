// Diamond-difference sweep for the zone-fastest nestings (DGZ, GDZ), where a
// zone row (fixed direction, group, j, k; i = 0..ni-1) is contiguous in HBM.
//
// Work split.  A row group of W lanes owns one k-row of the subdomain and
// walks its j-rows in sweep order, one row per step; lane l holds V
// consecutive zones of the row (W*V >= ni).  Row group kk runs kk steps behind
// row group 0 (wavefront over (j,k)); a unit strings R (direction, group)
// items back to back so the pipeline fill is paid once per unit.
//
// Row solve.  Along a row the incoming x-face obeys
//     f(i+1) = 2 psi_i - f(i),   psi_i = (q_i + a_i f(i)) / den_i,
// the affine map f -> alpha_i f + beta_i with alpha_i = 2 a_i/den_i - 1 and
// beta_i = 2 q_i/den_i (q_i: right-hand side plus the y and z face terms).
// Each lane composes its V maps, the row group scans the lane composites with
// DPP lane shifts (log2 W steps), and each lane then runs its V zones from the
// x-face the scan hands it.  The scan re-associates the recurrence: results
// differ from the sequential order by a few ulps per row.
//
// Faces.  The j-face stays in registers (same lane, next row); the k-face goes
// to row group kk+1 through LDS (double-buffered, one barrier per step).  All
// per-step HBM reads -- rhs, sigma_t, and the incoming x/j/k faces of the row
// -- are issued D steps ahead into a register ring, unconditionally (lanes
// that do not need a value read a hot address), so every step issues the
// same memory instructions and the loads of later steps stay in flight while
// a step computes.
#include "kr_launch.h"

namespace kr {

namespace {


typedef double d2 __attribute__((ext_vector_type(2)));
typedef unsigned v2u __attribute__((ext_vector_type(2)));
typedef unsigned v4u __attribute__((ext_vector_type(4)));

// Every HBM stream is a buffer resource (base, byte size): loads and stores
// take a 32-bit byte offset, and an offset past the end reads zeros / drops
// the store.  A lane that does not need a value (or a vacuum face, whose
// resource has size zero) uses NOWHERE, so every step issues the same memory
// instructions without branches.
typedef __amdgpu_buffer_rsrc_t Rsrc;
constexpr uint32_t NOWHERE = 0x80000000u;  // resources are smaller than 2 GiB
constexpr int RSRC_WORD3 = 0x00020000;      // raw 32-bit dword data, no swizzle

__device__ __forceinline__ Rsrc rsrc(const void *p, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)bytes, RSRC_WORD3);
}

constexpr int NT = 2;  // cache policy: non-temporal (streamed once)

// V consecutive doubles at byte offset off (16-byte aligned when V > 1).
template <int V, int AUX>
__device__ __forceinline__ void load_v(double (&t)[V], Rsrc r, uint32_t off) {
  if constexpr (V == 1) {
    t[0] = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, AUX));
  } else {
    static_assert(V == 2 || V == 4, "V is 1, 2 or 4");
#pragma unroll
    for (int h = 0; h < V / 2; ++h) {
      // (a NOWHERE offset stays out of range with the 16 bytes added)
      const d2 w = __builtin_bit_cast(d2, __builtin_amdgcn_raw_buffer_load_b128(r, off + 16u * h, 0, AUX));
      t[2 * h] = w.x;
      t[2 * h + 1] = w.y;
    }
  }
}

template <int V, int AUX>
__device__ __forceinline__ void store_v(Rsrc r, uint32_t off, const double (&t)[V]) {
  if constexpr (V == 1) {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(v2u, t[0]), r, off, 0, AUX);
  } else {
#pragma unroll
    for (int h = 0; h < V / 2; ++h) {
      d2 w;
      w.x = t[2 * h];
      w.y = t[2 * h + 1];
      __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(v4u, w), r, off + 16u * h, 0, AUX);
    }
  }
}

// Reciprocal: hardware estimate and two Newton steps (error about one ulp).
__device__ __forceinline__ double recip(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = __fma_rn(-x, r, 1.0);
  r = __fma_rn(r, e, r);
  e = __fma_rn(-x, r, 1.0);
  return __fma_rn(r, e, r);
}

// DPP move of a double; lanes without a source lane read zero.
template <int CTRL>
__device__ __forceinline__ double dpp(double v) {
  const long long s = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp((int)s, (int)s, CTRL, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_update_dpp((int)(s >> 32), (int)(s >> 32), CTRL, 0xF, 0xF, true);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

constexpr int DPP_ROW_SHR = 0x110;  // + shift (1..15), within 16-lane rows
constexpr int DPP_WAVE_SHR1 = 0x138;
constexpr int DPP_ROW_BCAST15 = 0x142;
constexpr int DPP_ROW_BCAST31 = 0x143;

// Two wait states between the VALU write of a value and a DPP read of it,
// as an explicit s_nop tied to the values (the asm "rewrites" them, so every
// DPP read of them comes after it) rather than whatever instructions the
// scheduler happens to place there.
__device__ __forceinline__ void dpp_wait(double &a, double &b) { asm volatile("s_nop 1" : "+v"(a), "+v"(b)); }
__device__ __forceinline__ void dpp_wait(double &a) { asm volatile("s_nop 1" : "+v"(a)); }

// DPP move of a double in which lanes without a source lane, and rows outside
// ROWS, keep `fill` (the move does not write them).
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_or(double v, double fill) {
  const long long s = __double_as_longlong(v), f = __double_as_longlong(fill);
  const int lo = __builtin_amdgcn_update_dpp((int)f, (int)s, CTRL, ROWS, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp((int)(f >> 32), (int)(s >> 32), CTRL, ROWS, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// One scan step: compose with the map CTRL brings in.  ID: a lane that gets
// no map composes with the identity (A = 1, B = 0) instead of zero.
template <int CTRL, int ROWS, bool ID>
__device__ __forceinline__ void compose_from(double &A, double &B) {
  dpp_wait(A, B);
  const double Ap = ID ? dpp_or<CTRL, ROWS>(A, 1.0) : dpp<CTRL>(A);
  const double Bp = ID ? dpp_or<CTRL, ROWS>(B, 0.0) : dpp<CTRL>(B);
  B = __fma_rn(A, Bp, B);
  A = A * Ap;
}

// Inclusive scan of affine maps f -> A f + B over the W lanes of a row group,
// lane 0 first.  Lane 0's map is constant (A = 0: it already holds the row's
// incoming face), so a lane whose partial composite reaches back to lane 0
// has A = 0 and is unchanged by whatever finite map it is composed with next.
// Within 16-lane rows (W <= 16) every lane's partial composite reaches lane 0
// by the time a shift runs past it, so out-of-row lanes may read zero.  A row
// group over several rows (W = 32, 64) starts rows 1.. mid-group: there a
// lane without a source must compose with the identity (with zeros, lane 16
// would drop the first 16 lanes' composite), and the cross-row broadcasts are
// masked as in a two-level scan (lane 15 into rows 1 and 3, then lane 31 into
// rows 2 and 3) so no segment is composed twice.
template <int W>
__device__ __forceinline__ void scan_maps(double &A, double &B) {
  constexpr bool ID = W > 16;
  if constexpr (W > 1) compose_from<DPP_ROW_SHR + 1, 0xF, ID>(A, B);
  if constexpr (W > 2) compose_from<DPP_ROW_SHR + 2, 0xF, ID>(A, B);
  if constexpr (W > 4) compose_from<DPP_ROW_SHR + 4, 0xF, ID>(A, B);
  if constexpr (W > 8) compose_from<DPP_ROW_SHR + 8, 0xF, ID>(A, B);
  if constexpr (W > 16) compose_from<DPP_ROW_BCAST15, 0xA, ID>(A, B);
  if constexpr (W > 32) compose_from<DPP_ROW_BCAST31, 0xC, ID>(A, B);
}

// Value of the lane one place back in the row group (lane 0: `first`).
template <int W>
__device__ __forceinline__ double lane_back(double v, double first, int l) {
  dpp_wait(v);
  const double u = W <= 16 ? dpp<DPP_ROW_SHR + 1>(v) : dpp<DPP_WAVE_SHR1>(v);
  return l == 0 ? first : u;
}

// Everything one block needs, read once from the tables.  In both
// zone-fastest nestings the (direction, group) item index p = pg_begin + r
// gives psi/rhs offset p*nz and face offsets p*nf (nf = faces per plane),
// since DGZ orders items d*ng + g and GDZ orders them g*nd + d.
struct RowArgs {
  const double *rhs, *sig, *coef;
  double *psi;
  const double *in[3];
  double *out[3];
  const double *w, *vol;  // quadrature weights, zone volumes (population)
  double vol0;            // the zone volume when all are equal (vol_uniform)
  int32_t vol_uniform;
  double *own2;           // the subdomain's own k-plane: faces between k-patches
  double *pop;            // population partial of this block, or nullptr
  int32_t nd, ng, ni, nj, nk, jd, kd;
  int32_t pg_begin, R, inner_is_g;
};

// Byte offsets of one row group's current row, stepped with additions.  A
// unit walks, per item, its k-patches (PK k-rows each, one per row group) and
// per patch the nj rows.  Each stream keeps one combined offset (item, k-row,
// j-row and the lane's first zone), so a cursor holds few registers.  Before
// the first row the cursor rests on row 0 of the first patch of the first
// item; past the last item it keeps the last item, so every offset stays in
// bounds.  Item offsets are relative to the unit's first item.
struct RowCursor {
  int32_t pos;      // rows since the unit's first row (negative: not started)
  int32_t jj, kp;   // row within the patch (sweep order), k-patch
  int32_t inner;    // inner pair index: g when the group is the inner index, else d
  int32_t bco;      // d*(ni+nj+nk+1): the item's coefficient record in LDS
  uint32_t zo8;     // rhs/psi: 8*(item*nz + (k*nj + j)*ni + first zone)
  uint32_t so8;     // sigma_t: 8*(g*nz + (k*nj + j)*ni + first zone)
  uint32_t vo8;     // zone volume: 8*((k*nj + j)*ni + first zone)
  uint32_t xo8;     // i-plane face: 8*(item*nf0 + k*nj + j)
  uint32_t yo8;     // j-plane face: 8*(item*nf1 + k*ni + first zone)
  uint32_t fo8;     // k-plane face: 8*(item*nf2 + j*ni + first zone)
};

struct RowConst {
  uint32_t nz8, nf8[3];
  int32_t L;
  int32_t jstep8, jrowstep8;       // per-row increments
  int32_t NPK;                     // k-patches per item
  int32_t dkz8, dkx8, dky8;        // offset steps from one k-patch to the next
  int32_t rows;                    // R*NPK*nj
};

// kz8 = 8*(k*nj*ni + first zone), kx8 = 8*k*nj, ky8 = 8*(k*ni + first zone)
// for the row group's k-row in the first patch; mb8 = 8*first zone.
__device__ __forceinline__ void cursor_init(RowCursor &c, const RowArgs &P, const RowConst &K, int pos,
                                            uint32_t kz8, uint32_t kx8, uint32_t ky8, uint32_t mb8) {
  c.pos = pos;
  c.jj = 0;
  c.kp = 0;
  const int j0 = P.jd > 0 ? 0 : P.nj - 1;
  const uint32_t j08 = 8u * (uint32_t)j0, jrow08 = j08 * (uint32_t)P.ni;
  const int p = P.pg_begin;
  int d, g;
  if (P.inner_is_g) {
    d = p / P.ng;
    g = p - d * P.ng;
    c.inner = g;
  } else {
    g = p / P.nd;
    d = p - g * P.nd;
    c.inner = d;
  }
  c.bco = d * K.L;
  c.zo8 = kz8 + jrow08;
  c.so8 = (uint32_t)g * K.nz8 + kz8 + jrow08;
  c.vo8 = kz8 + jrow08;
  c.xo8 = kx8 + j08;
  c.yo8 = ky8;
  c.fo8 = mb8 + jrow08;
}

__device__ __forceinline__ void cursor_step(RowCursor &c, const RowArgs &P, const RowConst &K) {
  if (++c.pos <= 0) return;
  if (++c.jj < P.nj) {
    c.zo8 += K.jrowstep8;
    c.so8 += K.jrowstep8;
    c.vo8 += K.jrowstep8;
    c.fo8 += K.jrowstep8;
    c.xo8 += K.jstep8;
    return;
  }
  // back to the patch's first row
  c.jj = 0;
  const int32_t jback = (P.nj - 1) * K.jrowstep8;
  c.zo8 -= jback;
  c.so8 -= jback;
  c.vo8 -= jback;
  c.fo8 -= jback;
  c.xo8 -= (P.nj - 1) * K.jstep8;
  if (c.pos >= K.rows) return;  // past the last item: stay on it
  if (++c.kp < K.NPK) {         // next k-patch of the same item
    c.zo8 += K.dkz8;
    c.so8 += K.dkz8;
    c.vo8 += K.dkz8;
    c.xo8 += K.dkx8;
    c.yo8 += K.dky8;
    return;
  }
  c.kp = 0;                     // first k-patch of the next item
  const int32_t kbz = (K.NPK - 1) * K.dkz8;
  c.zo8 += K.nz8 - kbz;
  c.so8 -= kbz;
  c.vo8 -= kbz;
  c.xo8 += K.nf8[0] - (K.NPK - 1) * K.dkx8;
  c.yo8 += K.nf8[1] - (K.NPK - 1) * K.dky8;
  c.fo8 += K.nf8[2];
  if (P.inner_is_g) {
    if (++c.inner == P.ng) {
      c.inner = 0;
      c.bco += K.L;
      c.so8 -= (P.ng - 1) * K.nz8;
    } else {
      c.so8 += K.nz8;
    }
  } else {
    if (++c.inner == P.nd) {
      c.inner = 0;
      c.bco = 0;
      c.so8 += K.nz8;
    } else {
      c.bco += K.L;
    }
  }
}

template <int SX, int V>
__device__ __forceinline__ constexpr int memo(int m) {
  return SX > 0 ? m : V - 1 - m;  // memory slot of sweep position m within a lane's block
}

template <int W, int V, int SX, bool POP, bool ADJ>
__device__ __forceinline__ void rows_body(const RowArgs &P, int PK, double *lds) {
  // steps of prefetch: 3 (2 with the population's extra registers)
  constexpr int RING = POP ? 2 : 3;
  const int nt = blockDim.x;
  const int tid = threadIdx.x;
  const int l = tid % W;
  const int kk = tid / W;
  const int ni = P.ni, nj = P.nj, nk = P.nk, nd = P.nd, ng = P.ng;
  const int nl = ni / V;                                // live lanes per row
  const bool live = l < nl;
  const int la = live ? l : 0;                          // address lane
  const int mb = SX > 0 ? la * V : ni - V - la * V;     // first zone (memory order) of the lane's block
  const int k = P.kd > 0 ? kk : nk - 1 - kk;  // k-row in the first k-patch

  const uint32_t nz = (uint32_t)ni * nj * nk;
  const uint32_t nf[3] = {(uint32_t)nj * nk, (uint32_t)ni * nk, (uint32_t)ni * nj};
  RowConst K;
  K.nz8 = 8 * nz;
#pragma unroll
  for (int x = 0; x < 3; ++x) K.nf8[x] = 8 * nf[x];
  K.L = ni + nj + nk + 1;  // record: A[ni], B[nj], C[nk], w
  K.jstep8 = P.jd > 0 ? 8 : -8;
  K.jrowstep8 = K.jstep8 * ni;
  K.NPK = nk / PK;
  K.dkz8 = (P.kd > 0 ? 8 : -8) * PK * nj * ni;
  K.dkx8 = (P.kd > 0 ? 8 : -8) * PK * nj;
  K.dky8 = (P.kd > 0 ? 8 : -8) * PK * ni;
  K.rows = P.R * K.NPK * nj;

  // LDS: k-face exchange [2][nt][V], then one coefficient record per
  // direction: {2cos_x/dx [ni], 2cos_y/dy [nj], 2cos_z/dz [nk], w}
  double *fzb = lds;
  double *co = lds + 2 * nt * V;
  {
    const gmem<double> *src = gp(P.coef);  // A[nd][ni], B[nd][nj], C[nd][nk], contiguous
    for (int e = tid; e < nd * (ni + nj + nk); e += nt) {
      int dst;
      if (e < nd * ni) {
        const int d = e / ni;
        dst = d * K.L + (e - d * ni);
      } else if (e < nd * (ni + nj)) {
        const int d = (e - nd * ni) / nj;
        dst = d * K.L + ni + (e - nd * ni - d * nj);
      } else {
        const int d = (e - nd * (ni + nj)) / nk;
        dst = d * K.L + ni + nj + (e - nd * (ni + nj) - d * nk);
      }
      co[dst] = src[e];
    }
    if (POP)
      for (int d = tid; d < nd; d += nt) co[d * K.L + K.L - 1] = gp(P.w)[d];
  }
  __syncthreads();

  // resources start at the unit's first item and cover its R items
  const int64_t p0 = P.pg_begin;
  const uint32_t span_z = (uint32_t)P.R * K.nz8;
  const Rsrc Rrhs = rsrc(P.rhs + p0 * nz, span_z);
  const Rsrc Rpsi = rsrc(P.psi + p0 * nz, span_z);
  const Rsrc Rsig = rsrc(P.sig, (uint32_t)ng * K.nz8);
  const Rsrc Rvol = rsrc(P.vol, POP ? K.nz8 : 0u);
  const uint32_t mb8 = 8u * (uint32_t)mb;
  Rsrc Rin[3], Rout[3];
#pragma unroll
  for (int x = 0; x < 3; ++x) {
    const uint32_t span_f = (uint32_t)P.R * K.nf8[x];
    Rin[x] = rsrc(P.in[x] ? P.in[x] + p0 * nf[x] : nullptr, P.in[x] ? span_f : 0u);  // vacuum: reads zeros
    Rout[x] = rsrc(P.out[x] + p0 * nf[x], span_f);
  }
  // k-faces between k-patches go through the own k-plane; the first patch's
  // incoming faces are there too when the subdomain has any (in[2] is the
  // own plane or absent)
  const bool has2 = P.in[2] != nullptr;
  const Rsrc Rown2 = rsrc(P.own2 + p0 * nf[2], (uint32_t)P.R * K.nf8[2]);

  double rq[RING][V], rs[RING][V], rz[RING][V], ry[RING][V], rx[RING];
  // zone volumes (population) come from L2 one step ahead; that load is
  // issued before the ring refill so waiting for it leaves the ring in flight
  double rv[V];

  // Each ring slot holds the cursor of the row it is loaded for, which is
  // the compute's cursor when the slot is consumed; a refill steps the
  // previous slot's cursor (the row before) into its own slot, so one cursor
  // update per step serves both the loads and the compute.
  RowCursor rc[RING];
  cursor_init(rc[0], P, K, -kk, 8u * (uint32_t)(k * nj * ni + mb), 8u * (uint32_t)(k * nj),
              8u * (uint32_t)(k * ni + mb), mb8);

  auto refill = [&](int r) {
    const RowCursor &lc = rc[r];
    load_v<V, NT>(rq[r], Rrhs, lc.zo8);
    load_v<V, 0>(rs[r], Rsig, lc.so8);
    // faces: the k-face only for row group 0, the j-face only on a patch's first row,
    // the x-face only for lane 0
    load_v<V, 0>(rz[r], Rown2, (kk == 0 && (lc.kp > 0 || has2)) ? lc.fo8 : NOWHERE);
    load_v<V, 0>(ry[r], Rin[1], lc.jj == 0 ? lc.yo8 : NOWHERE);
    {
      double t[1];
      load_v<1, 0>(t, Rin[0], l == 0 ? lc.xo8 : NOWHERE);
      rx[r] = t[0];
    }
  };

  // volumes: one value for the whole subdomain, or one L2 load per row
  auto vol_row = [&](uint32_t vo8) {
    if (P.vol_uniform) {
#pragma unroll
      for (int e = 0; e < V; ++e) rv[e] = P.vol0;
    } else {
      load_v<V, 0>(rv, Rvol, vo8);
    }
  };
  if (POP) vol_row(rc[0].vo8);
#pragma unroll
  for (int r = 0; r < RING; ++r) {
    if (r > 0) {
      rc[r] = rc[r - 1];
      cursor_step(rc[r], P, K);
    }
    refill(r);
  }

  double fy[V];
#pragma unroll
  for (int m = 0; m < V; ++m) fy[m] = 0.0;
  double pacc = 0.0;  // population: sum of (w*psi)*volume over this lane's zones
  const int T = K.rows + PK - 1;

  for (int t0 = 0; t0 < T; t0 += RING) {
#pragma unroll
    for (int r = 0; r < RING; ++r) {
      const int t = t0 + r;
      const int cur = t & 1, prev = cur ^ 1;
      const RowCursor &cc = rc[r];
      const bool act = cc.pos >= 0 && cc.pos < K.rows;
      const double bj = co[cc.bco + ni + (P.jd > 0 ? cc.jj : nj - 1 - cc.jj)];
      const int kc = P.kd > 0 ? cc.kp * PK + kk : nk - 1 - cc.kp * PK - kk;  // this row group's k-row
      const double ck = co[cc.bco + ni + nj + kc];
      const double bc = bj + ck;
      double a[V];
#pragma unroll
      for (int m = 0; m < V; ++m) a[m] = co[cc.bco + mb + memo<SX, V>(m)];

      // incoming faces of this row
      double fz[V];
      {
        const int src = kk == 0 ? tid : tid - W;
#pragma unroll
        for (int m = 0; m < V; ++m) {
          const double up = fzb[(prev * nt + src) * V + m];
          fz[m] = kk == 0 ? rz[r][memo<SX, V>(m)] : up;
        }
      }
      if (cc.jj == 0) {
#pragma unroll
        for (int m = 0; m < V; ++m) fy[m] = ry[r][memo<SX, V>(m)];
      }
      const double f0 = rx[r];

      // zone maps in sweep order, composed per lane; lane 0 absorbs the row's incoming face
      double inv[V], q[V];
      double Ac = 1.0, Bc = 0.0;
#pragma unroll
      for (int m = 0; m < V; ++m) {
        const int e = memo<SX, V>(m);
        const double den = (a[m] + bc) + rs[r][e];
        inv[m] = recip(den);
        // adjoint: incoming faces weigh 2 and outgoing faces a, b, c (the
        // transposed zone relations), so only q, B and the face updates change
        q[m] = ADJ ? __fma_rn(fz[m], 2.0, __fma_rn(fy[m], 2.0, rq[r][e]))
                   : __fma_rn(fz[m], ck, __fma_rn(fy[m], bj, rq[r][e]));
        const double i2 = inv[m] + inv[m];
        const double A = __fma_rn(a[m], i2, -1.0);
        const double B = ADJ ? (a[m] * q[m]) * inv[m] : q[m] * i2;
        if (m == 0) {
          Ac = A;
          Bc = B;
        } else {
          Bc = __fma_rn(A, Bc, B);
          Ac = A * Ac;
        }
      }
      if (l == 0) {
        Bc = __fma_rn(Ac, f0, Bc);
        Ac = 0.0;
      }
      scan_maps<W>(Ac, Bc);
      double f = lane_back<W>(Bc, f0, l);  // x-face entering the lane's first zone

      double p[V], fyo[V], fzo[V];
#pragma unroll
      for (int m = 0; m < V; ++m) {
        const int e = memo<SX, V>(m);
        if (ADJ) {
          p[e] = __fma_rn(2.0, f, q[m]) * inv[m];
          f = a[m] * p[e] - f;
          fyo[e] = bj * p[e] - fy[m];
          fzo[e] = ck * p[e] - fz[m];
        } else {
          p[e] = __fma_rn(a[m], f, q[m]) * inv[m];
          const double p2 = p[e] + p[e];
          f = p2 - f;
          fyo[e] = p2 - fy[m];
          fzo[e] = p2 - fz[m];
        }
        fy[m] = fyo[e];
      }

      const bool st = act && live;
      if (POP) {
        const double wd = co[cc.bco + K.L - 1];
        double t = 0.0;
#pragma unroll
        for (int e = 0; e < V; ++e) t += (wd * p[e]) * rv[e];
        pacc += st ? t : 0.0;
      }
      store_v<V, NT>(Rpsi, st ? cc.zo8 : NOWHERE, p);
      store_v<V, 0>(Rout[1], (st && cc.jj == nj - 1) ? cc.yo8 : NOWHERE, fyo);
      const bool ktop = st && kk == PK - 1, klast = cc.kp == K.NPK - 1;
      store_v<V, 0>(Rout[2], (ktop && klast) ? cc.fo8 : NOWHERE, fzo);
      store_v<V, 0>(Rown2, (ktop && !klast) ? cc.fo8 : NOWHERE, fzo);
      {
        const double t[1] = {f};
        store_v<1, 0>(Rout[0], (act && l == nl - 1) ? cc.xo8 : NOWHERE, t);
      }
#pragma unroll
      for (int m = 0; m < V; ++m) fzb[(cur * nt + tid) * V + m] = fzo[memo<SX, V>(m)];
      // the next step's row volumes (its slot was filled a step earlier)
      if (POP) vol_row(rc[(r + 1) % RING].vo8);
      {
        RowCursor nc = rc[(r + RING - 1) % RING];
        cursor_step(nc, P, K);
        rc[r] = nc;
      }
      refill(r);
      __syncthreads();
    }
  }
  if (POP) {
    // fixed-shape tree over the block: the partial does not depend on timing
    lds[tid] = pacc;
    __syncthreads();
    int half = 1;
    while (half * 2 < nt) half *= 2;
    for (; half > 0; half >>= 1) {
      if (tid < half && tid + half < nt) lds[tid] += lds[tid + half];
      __syncthreads();
    }
    if (tid == 0) P.pop[0] = lds[0];
  }
}

}  // namespace

// ADJ: the transposed sweep, psi-shaped rhsbar = H^-T psibar: it reads the
// psi field, writes the rhs field and runs every axis in reverse.
template <int W, int V, bool POP, bool ADJ>
__global__ __launch_bounds__(V == 4 ? 512 : 1024) void k_sweep_rows(const DevSdom *__restrict__ sd,
                                                     const SweepPlanes *__restrict__ planes,
                                                     const SweepUnit *__restrict__ units, SweepShape sh,
                                                     double *__restrict__ pop) {
  extern __shared__ double lds[];
  const SweepUnit U = units[blockIdx.x];
  const DevSdom &S = sd[U.sdom];
  const SweepPlanes &PL = planes[U.sdom];
  RowArgs P;
  P.rhs = ADJ ? S.psi : S.rhs;
  P.sig = S.sig;
  P.coef = S.coef[0];
  P.psi = ADJ ? S.rhs : S.psi;
#pragma unroll
  for (int x = 0; x < 3; ++x) {
    P.in[x] = PL.in[x];
    P.out[x] = PL.out[x];
  }
  P.nd = S.nd;
  P.ng = S.ng;
  P.ni = S.ni;
  P.nj = S.nj;
  P.nk = S.nk;
  P.jd = ADJ ? -S.jd : S.jd;
  P.kd = ADJ ? -S.kd : S.kd;
  P.pg_begin = U.pg_begin;
  P.R = U.pg_count;
  P.inner_is_g = sh.inner_is_g;
  P.w = S.w;
  P.vol = S.volume;
  P.vol_uniform = POP && S.vol_info[0] != 0.0;
  P.vol0 = POP ? S.vol_info[1] : 0.0;
  P.own2 = S.plane[2];
  P.pop = POP ? pop + blockIdx.x : nullptr;
  if ((S.id > 0) != ADJ)
    rows_body<W, V, 1, POP, ADJ>(P, sh.PK, lds);
  else
    rows_body<W, V, -1, POP, ADJ>(P, sh.PK, lds);
}

int launch_sweep_rows(hipStream_t st, const DevSdom *sd, const SweepPlanes *planes, const SweepUnit *units,
                      int32_t num_units, const SweepShape &shape, int64_t coef_doubles, double *pop, bool adjoint) {
  if (num_units <= 0) return KR_OK;
  const int W = shape.RG, V = shape.V;
  const int nt = W * shape.PK;
  const size_t lds = (2 * (size_t)nt * V + (size_t)coef_doubles) * sizeof(double);
  switch (W * 8 + V) {
#define X(w, v)                                                                                              \
  case w * 8 + v:                                                                                            \
    if (lds > 65536) {                                                                                       \
      int rc;                                                                                                \
      if ((rc = allow_lds((const void *)k_sweep_rows<w, v, false, true>, lds)) ||                            \
          (rc = allow_lds((const void *)k_sweep_rows<w, v, true, false>, lds)) ||                            \
          (rc = allow_lds((const void *)k_sweep_rows<w, v, false, false>, lds)))                             \
        return rc;                                                                                           \
    }                                                                                                        \
    if (adjoint)                                                                                             \
      hipLaunchKernelGGL((k_sweep_rows<w, v, false, true>), dim3(num_units), dim3(nt), lds, st, sd, planes,  \
                         units, shape, pop);                                                                 \
    else if (pop)                                                                                            \
      hipLaunchKernelGGL((k_sweep_rows<w, v, true, false>), dim3(num_units), dim3(nt), lds, st, sd, planes,  \
                         units, shape, pop);                                                                 \
    else                                                                                                     \
      hipLaunchKernelGGL((k_sweep_rows<w, v, false, false>), dim3(num_units), dim3(nt), lds, st, sd, planes, \
                         units, shape, pop);                                                                 \
    break;
    X(4, 1) X(8, 1) X(16, 1) X(32, 1) X(64, 1)
    X(4, 2) X(8, 2) X(16, 2) X(32, 2) X(64, 2)
    X(4, 4) X(8, 4) X(16, 4)
#undef X
    default: set_error("row sweep: unsupported row-group shape"); return KR_ERR_UNSUPPORTED;
  }
  KR_HIP_CHECK(hipGetLastError());
  return KR_OK;
}

}  // namespace kr
