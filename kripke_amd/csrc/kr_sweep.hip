// Diamond-difference sweep (src/Kripke/Kernel/SweepSubdomain.cpp:26-110)
// as a systolic pencil wavefront on gfx950, for the nestings whose fastest
// index is a group or a direction (DZG, ZDG, GZD, ZGD).
//
// A workgroup owns a patch of (j,k) zone columns of one subdomain, times Q
// adjacent (direction, group) pairs (adjacent in memory, so the Q lanes of a
// zone read one contiguous run).  Each thread walks its i-pencil one zone per
// step; the thread at (jj,kk) runs (jj+kk) steps behind the patch corner, so
// zone (i,j,k) is computed after its upwind neighbours (i-1,j,k), (i,j-1,k),
// (i,j,k-1).  The i-face stays in a register, the j- and k-faces pass to the
// neighbouring thread through LDS (double-buffered, one barrier per step).  A
// unit strings several pair groups and patches back to back through the same
// pipeline, so the wavefront fill is paid once per unit.  Faces that leave a
// patch go to the subdomain's own plane chunk (the reference's running-plane
// storage) and are read back by the next patch; faces that leave the
// subdomain go to the downwind subdomain's plane chunk (pipeline mode) or
// stay in the own chunk (Kernel::sweepSubdomain mode).
//
// Memory pipeline: rhs, sigma_t and the incoming faces of a step are read
// RING steps ahead into a register ring (buffer resources; a thread that
// needs no face reads NOWHERE, which returns zero without a memory access),
// and the step loop issues no other global load, so the wait before a
// step's use never drains the loads of later steps.  A patch-boundary face
// is written at least RING + 1 steps (and so at least one barrier) before
// the next patch's prefetch reads it; the shape chooser enforces that
// distance.
//
// Tables: the 2cos/delta rows and the quadrature weights sit in LDS, one
// record per direction: {a, b, c, w} when every row is uniform (the common
// case), whole rows otherwise, in two kernel instances that are both
// launched (each returns for the other's subdomains).
//
// k-ring variant (KRING): patches in k-fastest order, the faces between
// k-patches kept in an LDS ring instead of the own plane chunk.
//
// Arithmetic per zone (bit-identical across nestings and schedules):
//   num = fma(fz, c_k, fma(fy, b_j, fma(fx, a_i, rhs)))
//   psi = num / (((a_i + b_j) + c_k) + sigt)          (IEEE division)
//   f_out = 2*psi - f_in                              (exact scaling)
// with a_i = 2*xcos(d)/dx(i) etc. as the reference computes them per zone.
#include "kr_launch.h"

#include <algorithm>
#include <cstdlib>

namespace kr {

namespace {

constexpr int COL_RING = 2;  // steps of prefetch (each step more costs 10 VGPRs; at 4 steps, 98 VGPRs, the
                             // 64^3-zone 128-group 128-direction sweep ran 1-2 ms slower per iteration)

typedef __amdgpu_buffer_rsrc_t Rsrc;
constexpr uint32_t NOWHERE = 0x80000000u;  // resources are smaller than 2 GiB
constexpr int RSRC_WORD3 = 0x00020000;      // raw 32-bit dword data, no swizzle
constexpr int NT = 2;                       // cache policy: non-temporal

__device__ __forceinline__ Rsrc rsrc(const void *p, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)bytes, RSRC_WORD3);
}
template <int AUX>
__device__ __forceinline__ double ld(Rsrc r, uint32_t off) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, AUX));
}
template <int AUX>
__device__ __forceinline__ void st(Rsrc r, uint32_t off, double v) {
  typedef unsigned v2u __attribute__((ext_vector_type(2)));
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(v2u, v), r, off, 0, AUX);
}

struct ColArgs {
  const double *rhs, *sig, *coef, *w, *vol;
  const double *vol_info;  // {1 if every zone volume is the same, that volume} (k_vol_info)
  double *psi;
  const double *own[3];   // the subdomain's plane chunks
  double *out[3];         // destination of faces leaving the subdomain
  int32_t has_in[3];      // 0: vacuum face (reads as zero)
  double *pop;
  // strides (elements; each below 2^31) and the unit's first pair (d0, g0):
  // offsets are taken relative to that pair, where the resources are based
  int32_t fa, fg, fz, sg, sz;
  int32_t pa[3], pg[3], pf[3];
  int32_t d0, g0;
  int64_t base_f, base_p[3];  // element offset of (d0, g0) in psi/rhs and in the planes
  int32_t nd, ng, ni, nj, nk, jd, kd;
  int32_t pg_begin, R, inner_is_g;
};

// Position of one thread in its unit: u = ((r*NPK + pk)*NPJ + pj)*ni + ii
// (k-ring variant: ((r*NPJ + pj)*NPK + pk)*ni + ii),
// and the byte offsets of its current zone.  Before its first step a thread
// rests on u = 0; past the last pair group it keeps cycling the last one, so
// every offset stays in bounds.  Offsets move by additions only: one step
// along the pencil adds the per-step increments, the end of a pencil adds
// one row of the transition table (next j-patch, next k-patch, or back to
// the first patch), and a new pair group adds one of two pair rows (inner
// index advanced, or wrapped with the outer index advanced).
struct ColCursor {
  int32_t pos;               // steps since this thread's start (negative: not started)
  int32_t ii, pj, pk, r;
  uint32_t psi8, sig8;       // 8*(d*fa + g*fg + z*fz), 8*(g*sg + z*sz)
  uint32_t x8, y8, z8;       // face offsets: i-plane row (k,j); j-plane (k,i); k-plane (j,i)
  uint32_t vol8;             // 8*z
  int32_t ca, cb, cc, cw;    // LDS entries of a_i, b_j, c_k and w_d
};

struct ColConst {
  int Q, PJ, PK, NPJ, NPK;
  int q, jj, kk;
  int istep;                 // +1 / -1: zone i per step
  int gpo, r0;               // pair groups per outer index (inner extent / Q); the unit's first group mod gpo
  int rec;                   // LDS record per direction: COL_REC (uniform tables) or ni + nj + nk + 1
  int cstep;                 // step of the a_i entry along a pencil: 0 (uniform tables) or istep
  int32_t dpsi8, dsig8, dy8, dz8, dvol8;  // per-step offset increments
};

// Transition table (int32, in LDS): rows of COL_TAB_W entries.
//   rows 0-2, end of a pencil: next patch along the fast patch axis (j, or k
//     in k-fastest order) / next along the other axis / first patch again:
//     psi8, sig8, x8, y8, z8, vol8 increments and the changes of the a_i,
//     b_j, c_k entries in the LDS record (none with uniform tables)
//   rows 3-4, new pair group: inner index + Q / wrapped, outer index + 1:
//     psi8, sig8, x8, y8, z8 increments and the change of LDS record
//
// LDS record per direction: with every row of the 2cos/delta tables holding
// one value (the same zone width all along each axis, bit for bit: the
// generator's meshes) the kernel keeps {a, b, c, w}; otherwise the whole
// rows {A[ni], B[nj], C[nk], w} (kernel instance UNI = false).
constexpr int COL_TAB_W = 9;
constexpr int COL_REC = 4;
constexpr int COL_TAB_INTS = 5 * COL_TAB_W;
constexpr int COL_TAB_DOUBLES = 24;  // LDS reserved for the table (>= COL_TAB_INTS ints)
static_assert(2 * COL_TAB_DOUBLES >= COL_TAB_INTS, "transition table");

// KFAST: patches in k-fastest order (rows 0 and 1 move along k and j).
template <bool KFAST>
__device__ __forceinline__ int32_t col_entry(int e, const ColArgs &P, const ColConst &K) {
  const int row = e / COL_TAB_W, col = e - row * COL_TAB_W;
  const int jd = P.jd > 0 ? 1 : -1, kd = P.kd > 0 ? 1 : -1;
  int32_t v = 0;
  if (row < 3) {
    const int Di = -(P.ni - 1) * K.istep;
    const int jback = -jd * K.PJ * (K.NPJ - 1), kback = -kd * K.PK * (K.NPK - 1);
    int Dj, Dk;
    if (KFAST) {
      Dk = row == 0 ? kd * K.PK : kback;
      Dj = row == 0 ? 0 : (row == 1 ? jd * K.PJ : jback);
    } else {
      Dj = row == 0 ? jd * K.PJ : jback;
      Dk = row == 0 ? 0 : (row == 1 ? kd * K.PK : kback);
    }
    const int Dz = (Dk * P.nj + Dj) * P.ni + Di;
    switch (col) {
      case 0: v = 8 * Dz * P.fz; break;
      case 1: v = 8 * Dz * P.sz; break;
      case 2: v = 8 * (Dk * P.nj + Dj) * P.pf[0]; break;
      case 3: v = 8 * (Dk * P.ni + Di) * P.pf[1]; break;
      case 4: v = 8 * (Dj * P.ni + Di) * P.pf[2]; break;
      case 5: v = 8 * Dz; break;
      case 6: v = K.cstep != 0 ? Di : 0; break;
      case 7: v = K.cstep != 0 ? Dj : 0; break;
      default: v = K.cstep != 0 ? Dk : 0; break;
    }
  } else {
    int Ddd, Ddg;
    if (P.inner_is_g) {
      Ddd = row == 3 ? 0 : 1;
      Ddg = row == 3 ? K.Q : K.Q - P.ng;
    } else {
      Ddd = row == 3 ? K.Q : K.Q - P.nd;
      Ddg = row == 3 ? 0 : 1;
    }
    switch (col) {
      case 0: v = 8 * (Ddd * P.fa + Ddg * P.fg); break;
      case 1: v = 8 * Ddg * P.sg; break;
      case 2: v = 8 * (Ddd * P.pa[0] + Ddg * P.pg[0]); break;
      case 3: v = 8 * (Ddd * P.pa[1] + Ddg * P.pg[1]); break;
      case 4: v = 8 * (Ddd * P.pa[2] + Ddg * P.pg[2]); break;
      case 5: v = Ddd * K.rec; break;
      default: v = 0; break;
    }
  }
  return v;
}

// Every entry, whatever the workgroup size (a workgroup may have fewer
// threads than the table has entries).
template <bool KFAST>
__device__ __forceinline__ void col_table(int32_t *tab, const ColArgs &P, const ColConst &K) {
  for (int e = threadIdx.x; e < COL_TAB_INTS; e += blockDim.x) tab[e] = col_entry<KFAST>(e, P, K);
}

// Cursor at unit position 0 (the unit's first pair group, first patch, first zone).
__device__ __forceinline__ void col_init(ColCursor &c, const ColArgs &P, const ColConst &K, int pos) {
  c.pos = pos;
  c.ii = c.pj = c.pk = c.r = 0;
  const int pidx = P.pg_begin * K.Q + K.q;
  int d, g;
  if (P.inner_is_g) {
    d = pidx / P.ng;
    g = pidx - d * P.ng;
  } else {
    g = pidx / P.nd;
    d = pidx - g * P.nd;
  }
  const int j = P.jd > 0 ? K.jj : P.nj - 1 - K.jj, k = P.kd > 0 ? K.kk : P.nk - 1 - K.kk;
  const int i = K.istep > 0 ? 0 : P.ni - 1;
  // relative offsets stay below 2 GiB (host-checked unit spans): 32-bit math
  const uint32_t z = ((uint32_t)k * P.nj + j) * P.ni + i;
  const int dd = d - P.d0, dg = g - P.g0;
  c.psi8 = 8u * (uint32_t)(dd * P.fa + dg * P.fg + (int)z * P.fz);
  c.sig8 = 8u * (uint32_t)(g * P.sg + (int)z * P.sz);
  c.x8 = 8u * (uint32_t)(dd * P.pa[0] + dg * P.pg[0] + (k * P.nj + j) * P.pf[0]);
  c.y8 = 8u * (uint32_t)(dd * P.pa[1] + dg * P.pg[1] + (k * P.ni + i) * P.pf[1]);
  c.z8 = 8u * (uint32_t)(dd * P.pa[2] + dg * P.pg[2] + (j * P.ni + i) * P.pf[2]);
  c.vol8 = 8u * z;
  if (K.cstep != 0) {  // whole rows
    c.ca = d * K.rec + i;
    c.cb = d * K.rec + P.ni + j;
    c.cc = d * K.rec + P.ni + P.nj + k;
  } else {
    c.ca = d * K.rec;
    c.cb = d * K.rec + 1;
    c.cc = d * K.rec + 2;
  }
  c.cw = d * K.rec + K.rec - 1;
}

// COEF: also move the coefficient entries and the volume offset (the
// compute cursor; the prefetch cursor needs neither).  KFAST: patches in
// k-fastest order.  Every field is
// written once, after the branches, so that each stays a register (stores
// to different fields at the ends of different branches may otherwise be
// merged into one store through a selected address, which puts the cursor
// in scratch memory).
template <bool COEF, bool KFAST>
__device__ __forceinline__ void col_step(ColCursor &c, const ColArgs &P, const ColConst &K, const int32_t *tab) {
  if (++c.pos <= 0) return;
  int32_t dpsi = K.dpsi8, dsig = K.dsig8, dx = 0, dy = K.dy8, dz = K.dz8;
  int32_t dvol = K.dvol8, da = K.cstep, db = 0, dc = 0, dw = 0;
  int32_t ii = c.ii + 1, pj = c.pj, pk = c.pk, r = c.r;
  if (ii == P.ni) {  // end of a pencil
    ii = 0;
    int row = 0;
    if (KFAST) {
      if (++pk == K.NPK) {
        pk = 0;
        row = 1;
        if (++pj == K.NPJ) {
          pj = 0;
          row = 2;
        }
      }
    } else if (++pj == K.NPJ) {
      pj = 0;
      row = 1;
      if (++pk == K.NPK) {
        pk = 0;
        row = 2;
      }
    }
    const int32_t *t = tab + row * COL_TAB_W;
    dpsi = t[0];
    dsig = t[1];
    dx = t[2];
    dy = t[3];
    dz = t[4];
    dvol = t[5];
    da = t[6];
    db = t[7];
    dc = t[8];
    if (row == 2 && r + 1 < P.R) {  // past the last pair group: stay on it
      ++r;
      // the new group starts an outer index when its number is a multiple
      // of gpo (a remainder on the rare pair-group step instead of a
      // register per lane)
      const int32_t *u = tab + ((K.r0 + r) % K.gpo == 0 ? 4 : 3) * COL_TAB_W;
      dpsi += u[0];
      dsig += u[1];
      dx += u[2];
      dy += u[3];
      dz += u[4];
      da += u[5];
      db += u[5];
      dc += u[5];
      dw = u[5];
    }
  }
  c.ii = ii;
  c.pj = pj;
  c.pk = pk;
  c.r = r;
  c.psi8 += dpsi;
  c.sig8 += dsig;
  c.x8 += dx;
  c.y8 += dy;
  c.z8 += dz;
  if (COEF) {
    c.vol8 += dvol;
    c.ca += da;
    c.cb += db;
    c.cc += dc;
    c.cw += dw;
  }
}

// UNI: the instance for uniform tables (and, with the population, one
// volume for every zone); the other instance takes the whole table rows and
// per-zone volumes.  Both are launched for every stage; each unit's
// workgroup in the instance that does not match its subdomain returns at
// once (the flag and the volume check are read from the device).
template <bool POP, bool ADJ, bool KRING, bool UNI>
__device__ __forceinline__ void cols_body(const ColArgs &P, const ColConst &K, double *lds) {
  constexpr int RING = COL_RING;
  const int ni = P.ni, nj = P.nj, nk = P.nk, nd = P.nd, ng = P.ng;
  const int ncoef = nd * (ni + nj + nk);
  const gmem<double> *src = gp(P.coef);  // A[nd][ni], B[nd][nj], C[nd][nk], then the uniform flag
  {
    const bool fast = src[ncoef] != 0.0 && (!POP || gp(P.vol_info)[0] != 0.0);
    if (fast != UNI) return;  // the whole workgroup: its subdomain is the other instance's
  }
  const int nt = blockDim.x;
  const int tid = threadIdx.x;
  const int L = K.rec;
  int32_t *tab = (int32_t *)lds;  // the cursor transition table
  double *fyb = lds + COL_TAB_DOUBLES;  // [2][nt]
  double *fzb = fyb + 2 * nt;           // [2][nt]
  double *co = fyb + 4 * nt;            // [nd][L]
  // KRING (patches in k-fastest order): the faces between k-patches,
  // [ni][PJ][Q] by the zone's step along its pencil; the top layer of patch
  // (pj, pk) writes a slot ni - PK + 1 steps before the bottom layer of the
  // next patch, (pj, pk + 1), reads it, and rewrites it only after that read
  double *kring = co + nd * L;
  if (UNI) {
    for (int d = tid; d < nd; d += nt) {
      co[d * L + 0] = src[d * ni];
      co[d * L + 1] = src[nd * ni + d * nj];
      co[d * L + 2] = src[nd * (ni + nj) + d * nk];
    }
  } else {
    for (int e = tid; e < ncoef; e += nt) {
      int dst;
      if (e < nd * ni) {
        const int d = e / ni;
        dst = d * L + (e - d * ni);
      } else if (e < nd * (ni + nj)) {
        const int d = (e - nd * ni) / nj;
        dst = d * L + ni + (e - nd * ni - d * nj);
      } else {
        const int d = (e - nd * (ni + nj)) / nk;
        dst = d * L + ni + nj + (e - nd * (ni + nj) - d * nk);
      }
      co[dst] = src[e];
    }
  }
  for (int d = tid; d < nd; d += nt) co[d * L + L - 1] = POP ? gp(P.w)[d] : 0.0;
  col_table<KRING>(tab, P, K);
  __syncthreads();

  // resources start at the unit's first pair; the host keeps each unit's
  // span below 2 GiB (unit_spans_fit), so 32-bit offsets cover it
  const int64_t items = (int64_t)nd * ng;
  const uint32_t nz = (uint32_t)ni * nj * nk;
  const int64_t nf[3] = {(int64_t)nj * nk, (int64_t)ni * nk, (int64_t)ni * nj};
  auto span = [](int64_t elems) { return (uint32_t)std::min<int64_t>(elems * 8, 0x7FFFFFF8LL); };
  const Rsrc Rrhs = rsrc(P.rhs + P.base_f, span(items * nz - P.base_f));
  const Rsrc Rpsi = rsrc(P.psi + P.base_f, span(items * nz - P.base_f));
  const Rsrc Rsig = rsrc(P.sig, span((int64_t)ng * nz));
  // the tables and the zone volumes take plain global loads (always in
  // range, two scalar registers of base each instead of a four-register
  // resource: the kernel is short of scalar registers)
  const gmem<double> *gvol = gp(P.vol);
  // x-faces (one load at a pencil's start, one store at its end, both under
  // a branch): plain global accesses from scalar bases
  const gmem<char> *gin0 = (const gmem<char> *)gp(P.own[0] + P.base_p[0]);
  gmem<char> *gout0 = (gmem<char> *)gp(P.out[0] + P.base_p[0]);
  const bool in0 = P.has_in[0] != 0;
  const Rsrc Rown1 = rsrc(P.own[1] + P.base_p[1], span(items * nf[1] - P.base_p[1]));
  const Rsrc Rown2 = rsrc(P.own[2] + P.base_p[2], span(items * nf[2] - P.base_p[2]));
  const Rsrc Rout1 = rsrc(P.out[1] + P.base_p[1], span(items * nf[1] - P.base_p[1]));
  const Rsrc Rout2 = rsrc(P.out[2] + P.base_p[2], span(items * nf[2] - P.base_p[2]));

  double rq[RING], rs[RING], rx[RING], ry[RING], rz[RING];
#pragma unroll
  for (int r = 0; r < RING; ++r) rx[r] = ry[r] = rz[r] = 0.0;
  ColCursor lc, cc;
  col_init(lc, P, K, -(K.jj + K.kk));
  cc = lc;

  auto refill = [&](int r) {
    rq[r] = ld<NT>(Rrhs, lc.psi8);
    rs[r] = ld<0>(Rsig, lc.sig8);
    // incoming faces: x at a pencil start; y for the first j-row of a patch
    // (the subdomain's incoming face on the first patch, else the previous
    // patch's outgoing face in the own chunk); z likewise
    // (each ring slot written on every path: conditional stores to different
    // slots may be merged into one through a selected address, which puts
    // the ring in scratch memory)
    double vx = rx[r], vy = ry[r], vz = rz[r];
    if (lc.ii == 0) vx = in0 ? *(const gmem<double> *)(gin0 + lc.x8) : 0.0;
    if (K.jj == 0) vy = ld<0>(Rown1, (lc.pj > 0 || P.has_in[1]) ? lc.y8 : NOWHERE);
    if (K.kk == 0) vz = ld<0>(Rown2, ((!KRING && lc.pk > 0) || (lc.pk == 0 && P.has_in[2])) ? lc.z8 : NOWHERE);
    rx[r] = vx;
    ry[r] = vy;
    rz[r] = vz;
    col_step<false, KRING>(lc, P, K, tab);
  };

  // zone volumes: one value for the whole subdomain (the UNI instance), else
  // one load per zone step, issued a step ahead and before that step's ring
  // loads (so its wait does not wait for them)
  // (a global-space read: a generic one would leave a flat access pending at
  // the loop head, and every wait in the loop would then drain the ring)
  const double vol0 = POP ? gp(P.vol_info)[1] : 0.0;
  double rv = vol0;
  if (POP && !UNI) rv = gvol[cc.vol8 / 8u];
#pragma unroll
  for (int r = 0; r < RING; ++r) refill(r);

  const int rows = P.R * K.NPK * K.NPJ * ni;
  const int T = rows + (K.PJ - 1) + (K.PK - 1);
  double fx = 0.0, pacc = 0.0;
  // LDS slots of the upwind neighbours' faces, read on every step and used
  // where the thread has such a neighbour (a thread on the patch's first row
  // reads its own slot and takes the prefetched face instead)
  const int iy = K.jj == 0 ? tid : tid - K.Q, iz = K.kk == 0 ? tid : tid - K.Q * K.PJ;

  for (int t0 = 0; t0 < T; t0 += RING) {
#pragma unroll
    for (int r = 0; r < RING; ++r) {
      const int cur = (t0 + r) & 1, prev = cur ^ 1;
      const bool act = cc.pos >= 0 && cc.pos < rows;
      const double a = co[cc.ca];
      const double bj = co[cc.cb];
      const double ck = co[cc.cc];
      if (cc.ii == 0) fx = rx[r];
      const double ly = fyb[prev * nt + iy], lz = fzb[prev * nt + iz];
      const double fy = K.jj == 0 ? ry[r] : ly;
      double fzv = K.kk == 0 ? rz[r] : lz;
      if (KRING) {
        const double lr = kring[(cc.ii * K.PJ + K.jj) * K.Q + K.q];
        if (K.kk == 0 && cc.pk > 0) fzv = lr;
      }
      // adjoint: incoming faces weigh 2, outgoing faces a, b, c (transposed zone relations)
      const double wx = ADJ ? 2.0 : a, wy = ADJ ? 2.0 : bj, wz = ADJ ? 2.0 : ck;
      const double num = __fma_rn(fzv, wz, __fma_rn(fy, wy, __fma_rn(fx, wx, rq[r])));
      const double den = ((a + bj) + ck) + rs[r];
      const double p = num / den;
      const double p2 = 2.0 * p;
      fx = (ADJ ? a * p : p2) - fx;
      const double fyo = (ADJ ? bj * p : p2) - fy, fzo = (ADJ ? ck * p : p2) - fzv;
      st<NT>(Rpsi, act ? cc.psi8 : NOWHERE, p);
      if (act && cc.ii == ni - 1) *(gmem<double> *)(gout0 + cc.x8) = fx;
      // both stores of a face issue with the unused one aimed NOWHERE: the
      // lanes of a wave disagree on which buffer a face goes to, and a store
      // through a per-lane choice of resource would loop over the choices
      if (act && K.jj == K.PJ - 1) {
        const bool last = cc.pj == K.NPJ - 1;
        st<0>(Rout1, last ? cc.y8 : NOWHERE, fyo);
        st<0>(Rown1, last ? NOWHERE : cc.y8, fyo);
      }
      if (act && K.kk == K.PK - 1) {
        const bool last = cc.pk == K.NPK - 1;
        if (KRING) {
          if (last)
            st<0>(Rout2, cc.z8, fzo);
          else
            kring[(cc.ii * K.PJ + K.jj) * K.Q + K.q] = fzo;
        } else {
          st<0>(Rout2, last ? cc.z8 : NOWHERE, fzo);
          st<0>(Rown2, last ? NOWHERE : cc.z8, fzo);
        }
      }
      fyb[cur * nt + tid] = fyo;
      fzb[cur * nt + tid] = fzo;
      if (POP) {
        const double t = (co[cc.cw] * p) * rv;
        pacc += act ? t : 0.0;
        asm volatile("" : "+v"(pacc));  // summed in its step, not deferred
      }
      col_step<true, KRING>(cc, P, K, tab);
      if (POP && !UNI) rv = gvol[cc.vol8 / 8u];
      refill(r);
      __syncthreads();
    }
  }
  if (POP) {
    lds[tid] = pacc;  // the face buffers are free after the last barrier
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

// ADJ: the transposed sweep (reads psi, writes rhs, every axis reversed).
template <bool POP, bool ADJ, bool KRING, bool UNI>
__device__ __forceinline__ void sweep_unit(const DevSdom *__restrict__ sd, const SweepPlanes *__restrict__ planes,
                                           const SweepUnit *__restrict__ units, const SweepShape &sh,
                                           double *__restrict__ pop) {
  extern __shared__ double lds[];
  const SweepUnit U = units[blockIdx.x];
  const DevSdom &S = sd[U.sdom];
  const SweepPlanes &PL = planes[U.sdom];
  ColArgs P;
  P.rhs = ADJ ? S.psi : S.rhs;
  P.sig = S.sig;
  P.coef = S.coef[0];
  P.w = S.w;
  P.vol = S.volume;
  P.vol_info = S.vol_info;
  P.psi = ADJ ? S.rhs : S.psi;
#pragma unroll
  for (int x = 0; x < 3; ++x) {
    P.own[x] = S.plane[x];
    P.out[x] = PL.out[x];
    P.has_in[x] = PL.in[x] != nullptr;
    P.pa[x] = (int32_t)S.pa[x];
    P.pg[x] = (int32_t)S.pg[x];
    P.pf[x] = (int32_t)S.pf[x];
  }
  P.pop = POP ? pop + blockIdx.x : nullptr;
  P.fa = (int32_t)S.fa;
  P.fg = (int32_t)S.fg;
  P.fz = (int32_t)S.fz;
  P.sg = (int32_t)S.sg;
  P.sz = (int32_t)S.sz;
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
  {
    const int pidx0 = U.pg_begin * sh.Q;  // the unit's first pair: smallest offsets in every field
    P.d0 = sh.inner_is_g ? pidx0 / S.ng : pidx0 % S.nd;
    P.g0 = sh.inner_is_g ? pidx0 % S.ng : pidx0 / S.nd;
    P.base_f = P.d0 * S.fa + P.g0 * S.fg;
#pragma unroll
    for (int x = 0; x < 3; ++x) P.base_p[x] = P.d0 * S.pa[x] + P.g0 * S.pg[x];
  }
  ColConst K;
  K.Q = sh.Q;
  K.PJ = sh.PJ;
  K.PK = sh.PK;
  K.NPJ = sh.NPJ;
  K.NPK = sh.NPK;
  const int tid = threadIdx.x;
  K.q = tid % sh.Q;
  K.jj = (tid / sh.Q) % sh.PJ;
  K.kk = tid / (sh.Q * sh.PJ);
  K.istep = ((S.id > 0) != ADJ) ? 1 : -1;
  K.gpo = (sh.inner_is_g ? S.ng : S.nd) / sh.Q;
  K.r0 = U.pg_begin % K.gpo;
  K.dpsi8 = (int32_t)(8 * K.istep * S.fz);
  K.dsig8 = (int32_t)(8 * K.istep * S.sz);
  K.dy8 = (int32_t)(8 * K.istep * S.pf[1]);
  K.dz8 = (int32_t)(8 * K.istep * S.pf[2]);
  K.dvol8 = 8 * K.istep;
  K.rec = UNI ? COL_REC : S.ni + S.nj + S.nk + 1;
  K.cstep = UNI ? 0 : K.istep;
  cols_body<POP, ADJ, KRING, UNI>(P, K, lds);
}

// Held to 80 VGPRs (six waves per SIMD), so that three 512-thread
// workgroups share a CU.
template <bool ADJ, bool KRING, bool UNI>
__global__ __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(6))) void k_sweep(
    const DevSdom *__restrict__ sd, const SweepPlanes *__restrict__ planes, const SweepUnit *__restrict__ units,
    SweepShape sh, double *__restrict__ pop) {
  sweep_unit<false, ADJ, KRING, UNI>(sd, planes, units, sh, pop);
}

// The solve's sweep, with the fused population.
template <bool KRING, bool UNI>
__global__ __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(6))) void k_sweep_pop(
    const DevSdom *__restrict__ sd, const SweepPlanes *__restrict__ planes, const SweepUnit *__restrict__ units,
    SweepShape sh, double *__restrict__ pop) {
  sweep_unit<true, false, KRING, UNI>(sd, planes, units, sh, pop);
}

// 2*cos(d)/delta per (direction, zone index), the exact expression of
// SweepSubdomain.cpp:79-81 (product by 2 is exact, IEEE division), for the
// subdomains [first, first+count).  Recomputed before every sweep so that
// the quadrature and mesh fields are read as they are at call time.
__global__ __launch_bounds__(256) void k_sweep_coef(const DevSdom *__restrict__ sd, int first) {
  const DevSdom &S = sd[first + blockIdx.x];
  const int n[3] = {S.ni, S.nj, S.nk};
  const int total = S.nd * (n[0] + n[1] + n[2]);
  int same = 1;  // every row holds one value (bit for bit)
  for (int e = threadIdx.x; e < total; e += blockDim.x) {
    const int d = e / (n[0] + n[1] + n[2]);
    int r = e - d * (n[0] + n[1] + n[2]);
    int ax = 0;
    while (r >= n[ax]) {
      r -= n[ax];
      ++ax;
    }
    const double c2 = 2.0 * gp(S.cosines[ax])[d];
    const double v = c2 / gp(S.delta[ax])[r];
    gp(S.coef[ax])[d * n[ax] + r] = v;
    same &= __double_as_longlong(v) == __double_as_longlong(c2 / gp(S.delta[ax])[0]);
  }
  same = __syncthreads_and(same);
  // the flag after the tables: which instance of the systolic sweep takes the subdomain
  if (threadIdx.x == 0) gp(S.coef[0])[total] = same ? 1.0 : 0.0;
}

int allow_lds(const void *kernel, size_t bytes) {
  KR_HIP_CHECK(hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
  return KR_OK;
}

int launch_sweep_coef(hipStream_t st, const DevSdom *sd, int first, int count) {
  if (count <= 0) return KR_OK;
  hipLaunchKernelGGL(k_sweep_coef, dim3(count), dim3(256), 0, st, sd, first);
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
// k-row.  Two zones per lane by default; four (rows of up to 64 zones, at
// most 512 threads: that kernel's launch bound) only when v_max asks for
// them.  Returns false when the subdomain does not fit one workgroup.
static bool row_shape(SweepShape &sh, int32_t nd, int32_t ng, int32_t ni, int32_t nj, int32_t nk, bool aligned16,
                      int v_max, size_t lds_limit) {
  int want = ni % 4 == 0 ? 4 : ni % 2 == 0 ? 2 : 1;
  want = v_max > 0 ? std::min(want, v_max) : std::min(want, 2);
  const int64_t nz = (int64_t)ni * nj * nk;
  const int64_t maxface = std::max({(int64_t)nj * nk, (int64_t)ni * nk, (int64_t)ni * nj});
  // every stream is a buffer resource based at the unit's first item and
  // addressed with 32-bit byte offsets below 2 GiB (units hold <= 64 items)
  if (64 * nz * 8 >= (1LL << 31) || 64 * maxface * 8 >= (1LL << 31) || (int64_t)ng * nz * 8 >= (1LL << 31))
    return false;
  for (int V = 4; V >= 1; V /= 2) {
    if (V > want || ni % V != 0 || (V > 1 && !aligned16)) continue;
    int W = 4;
    while (W * V < ni) W *= 2;
    if (W > (V == 4 ? 16 : 64)) continue;
    const int tmax = V == 4 ? 512 : 1024;
    // k-rows per block: all of them, or a divisor of nk; between k-patches the
    // k-face goes through the own plane and must be written (by the patch's
    // top row group, PK-1 steps into a row) before the next patch prefetches
    // it (4 steps ahead): nj >= PK + 4
    for (int PK = largest_divisor_le(nk, tmax / W); PK >= 1; PK = largest_divisor_le(nk, PK - 1)) {
      const size_t lds = (2 * (size_t)W * PK * V + (size_t)nd * (ni + nj + nk + 1)) * sizeof(double);
      if ((PK == nk || nj >= PK + 4) && lds <= lds_limit) {
        sh.kind = 1;
        sh.RG = W;
        sh.V = V;
        sh.Q = 1;
        sh.PJ = nj;
        sh.PK = PK;
        sh.NPJ = 1;
        sh.NPK = nk / PK;
        return true;
      }
      if (PK == 1) break;
    }
  }
  return false;
}

SweepShape choose_sweep_shape(int nesting, int32_t nd, int32_t ng, int32_t ni, int32_t nj, int32_t nk, int ref_mode,
                              bool aligned16, const SweepTune &tune) {
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
  // prefer shapes within 64 KiB of LDS (two workgroups per CU), then the device maximum
  if (zfast && (row_shape(sh, nd, ng, ni, nj, nk, aligned16, tune.v_max, 65536) ||
                (tune.lds_max > 65536 && row_shape(sh, nd, ng, ni, nj, nk, aligned16, tune.v_max, tune.lds_max))))
    return sh;
  if (!zfast && tune.v_max == 2) {
    const int kt_max = 16 / std::max(1, tune.cap_div);
    if (waves_shape(sh, nd, ng, ni, nj, nk, std::min<size_t>(65536, tune.lds_max), kt_max) ||
        (tune.lds_max > 65536 && waves_shape(sh, nd, ng, ni, nj, nk, tune.lds_max, kt_max)))
      return sh;
  }
  if (zfast) {
    sh.Q = 1;
  } else {
    int qv = 1;
    while (qv < 16 && inner % (qv * 2) == 0) qv *= 2;
    sh.Q = qv;
  }
  // systolic kernel with its patches in k-fastest order (tuning point
  // v_max = 3): the faces between k-patches stay in an LDS ring of
  // [ni][PJ][Q] doubles, and only the faces between j-patches go through the
  // own plane chunk (16/PJ bytes per zone and pair).  Among the shapes with
  // the most columns within the thread budget (1024 / Q / cap_div), the
  // widest PJ whose LDS fits: 64 KiB first (two workgroups per CU), then the
  // device maximum.  Distances: a ring slot is written ni - PK + 1 steps
  // before the next k-patch reads it (ni >= PK); a j-face leaves NPK*ni -
  // PJ + 1 steps before the next j-patch uses it, and is read COL_RING steps
  // ahead of that (NPK*ni >= PJ + COL_RING).
  if (!zfast && tune.v_max == 3) {
    const size_t coef = (size_t)nd * COL_REC;
    // the instance for varying tables holds whole table rows: it must fit the device too
    const size_t rows_extra = (size_t)nd * (ni + nj + nk + 1 - COL_REC) * sizeof(double);
    const int cap = 1024 / sh.Q / std::max(1, tune.cap_div);
    for (int pass = 0; pass < 2; ++pass) {
      const size_t lim = pass == 0 ? std::min<size_t>(65536, tune.lds_max) : tune.lds_max;
      int best_pj = 0, best_pk = 0;
      for (int pj = largest_divisor_le(nj, cap); pj >= 1; pj = largest_divisor_le(nj, pj - 1)) {
        for (int pk = largest_divisor_le(nk, cap / pj); pk >= 1; pk = largest_divisor_le(nk, pk - 1)) {
          const size_t nt = (size_t)sh.Q * pj * pk;
          const size_t lds = (COL_TAB_DOUBLES + 4 * nt + coef + (size_t)ni * pj * sh.Q) * sizeof(double);
          const bool ok = ni >= pk && (nj == pj || (int64_t)(nk / pk) * ni >= pj + COL_RING) && lds <= lim &&
                          lds + rows_extra <= tune.lds_max;
          if (ok && (pj * pk > best_pj * best_pk || (pj * pk == best_pj * best_pk && pj > best_pj))) {
            best_pj = pj;
            best_pk = pk;
          }
          if (pk == 1) break;
        }
        if (pj == 1) break;
      }
      if (best_pj > 0) {
        sh.PJ = best_pj;
        sh.PK = best_pk;
        sh.NPJ = nj / best_pj;
        sh.NPK = nk / best_pk;
        sh.kring = sh.NPK > 1 ? 1 : 0;
        return sh;
      }
    }
  }
  // LDS: j/k face exchange [4][threads] and the direction records [nd][COL_REC]
  const size_t coef_lds = (size_t)nd * COL_REC * sizeof(double);
  const size_t rows_extra = (size_t)nd * (ni + nj + nk + 1 - COL_REC) * sizeof(double);
  for (int pass = 0; pass < 2; ++pass) {
  const size_t lds_limit = pass == 0 ? std::min<size_t>(65536, tune.lds_max) : tune.lds_max;
  for (int cap = 1024 / sh.Q / std::max(1, tune.cap_div); cap >= 1; cap /= 2) {
    // patch: as many (j,k) columns as fit, subject to the pipeline distances
    // (a patch's downstream neighbour prefetches a face COL_RING steps before
    // it uses it, and must do so at least one step after the face is written).
    // Among patches of that many columns, the one with the least face traffic
    // between patches: a face crossing between patches is written to and read
    // back from the own plane chunk, 16 bytes per pair per patch extent along
    // that axis, so the cost is 1/pj + 1/pk over the axes cut into patches
    // (64 columns of 32x32: 8x8 moves 4 B per zone-pair, 32x2 moves 8.5 B).
    int best_pj = 0, best_pk = 0;
    double best_cost = 0.0;
    for (int pj = largest_divisor_le(nj, cap); pj >= 1; pj = largest_divisor_le(nj, pj - 1)) {
      for (int pk = largest_divisor_le(nk, cap / pj); pk >= 1; pk = largest_divisor_le(nk, pk - 1)) {
        const int npj = nj / pj, npk = nk / pk;
        const bool ok_j = (npj == 1) || (ni >= pj + COL_RING);
        const bool ok_k = (npk == 1) || ((int64_t)npj * ni >= pk + COL_RING);
        const double cost = (npj > 1 ? 1.0 / pj : 0.0) + (npk > 1 ? 1.0 / pk : 0.0);
        if (ok_j && ok_k &&
            (pj * pk > best_pj * best_pk || (pj * pk == best_pj * best_pk && cost < best_cost))) {
          best_pj = pj;
          best_pk = pk;
          best_cost = cost;
        }
        if (pk == 1) break;
      }
      if (pj == 1) break;
    }
    if (best_pj == 0) continue;
    const size_t lds = coef_lds + (COL_TAB_DOUBLES + 4 * (size_t)sh.Q * best_pj * best_pk) * sizeof(double);
    if (lds > lds_limit || lds + rows_extra > tune.lds_max) continue;
    sh.PJ = best_pj;
    sh.PK = best_pk;
    sh.NPJ = nj / best_pj;
    sh.NPK = nk / best_pk;
    return sh;
  }
  }
  sh.kind = 2;  // no shape fits the device's LDS
  return sh;
}

// LDS of one systolic workgroup: uniform tables (one record per direction)
// or whole table rows.
static size_t cols_lds_bytes(const SweepShape &shape, bool uni) {
  const size_t nt = (size_t)shape.Q * shape.PJ * shape.PK;
  const size_t ring = shape.kring ? (size_t)shape.ni * shape.PJ * shape.Q : 0;
  const size_t rec = uni ? COL_REC : (size_t)shape.ni + shape.nj + shape.nk + 1;
  return (COL_TAB_DOUBLES + 4 * nt + (size_t)shape.nd * rec + ring) * sizeof(double);
}

size_t sweep_lds_bytes(const SweepShape &shape, int64_t coef_doubles) {
  if (shape.kind == 1) return (2 * (size_t)shape.RG * shape.PK * shape.V + (size_t)coef_doubles) * sizeof(double);
  const size_t nt = (size_t)shape.Q * shape.PJ * shape.PK;
  if (shape.kind == 3) return std::max(2 * nt + (size_t)coef_doubles, nt) * sizeof(double);
  return cols_lds_bytes(shape, false);  // the larger of the systolic kernel's two instances
}

int launch_sweep(hipStream_t st, const DevSdom *sd, const SweepPlanes *planes, const SweepUnit *units,
                 int32_t num_units, const SweepShape &shape, double *pop, bool adjoint) {
  if (num_units <= 0) return KR_OK;
  if (adjoint && pop) {
    set_error("sweep: no population in the adjoint sweep");
    return KR_ERR_INVALID;
  }
  if (shape.kind == 3) return launch_sweep_waves(st, sd, planes, units, num_units, shape, pop, adjoint);
  if (shape.kind == 1)
    return launch_sweep_rows(st, sd, planes, units, num_units, shape,
                             (int64_t)shape.nd * (shape.ni + shape.nj + shape.nk + 1), pop, adjoint);
  const int nt = shape.Q * shape.PJ * shape.PK;
  // both instances: each unit's workgroup in the one that does not match
  // its subdomain's tables (uniform or not) returns at once
#define KR_SWEEP_LAUNCH(KR, UNI)                                                                                   \
  do {                                                                                                             \
    const size_t lds = cols_lds_bytes(shape, UNI);                                                                 \
    if (lds > 65536) {                                                                                             \
      int rc;                                                                                                      \
      if ((rc = allow_lds((const void *)k_sweep<true, KR, UNI>, lds)) ||                                           \
          (rc = allow_lds((const void *)k_sweep_pop<KR, UNI>, lds)) ||                                             \
          (rc = allow_lds((const void *)k_sweep<false, KR, UNI>, lds)))                                            \
        return rc;                                                                                                 \
    }                                                                                                              \
    if (adjoint)                                                                                                   \
      hipLaunchKernelGGL((k_sweep<true, KR, UNI>), dim3(num_units), dim3(nt), lds, st, sd, planes, units, shape,   \
                         pop);                                                                                     \
    else if (pop)                                                                                                  \
      hipLaunchKernelGGL((k_sweep_pop<KR, UNI>), dim3(num_units), dim3(nt), lds, st, sd, planes, units, shape,     \
                         pop);                                                                                     \
    else                                                                                                           \
      hipLaunchKernelGGL((k_sweep<false, KR, UNI>), dim3(num_units), dim3(nt), lds, st, sd, planes, units,        \
                         shape, pop);                                                                              \
    KR_HIP_CHECK(hipGetLastError());                                                                               \
  } while (0)
  if (shape.kring) {
    KR_SWEEP_LAUNCH(true, true);
    KR_SWEEP_LAUNCH(true, false);
  } else {
    KR_SWEEP_LAUNCH(false, true);
    KR_SWEEP_LAUNCH(false, false);
  }
#undef KR_SWEEP_LAUNCH
  return KR_OK;
}

}  // namespace kr
