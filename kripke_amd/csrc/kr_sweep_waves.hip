// Diamond-difference sweep (src/Kripke/Kernel/SweepSubdomain.cpp:26-110) for
// the nestings whose fastest index is a group or a direction (DZG, ZDG, GZD,
// ZGD), with wave-uniform control.
//
// Lanes.  The 64 lanes of a wave take 64 (direction, group) pairs that are
// adjacent in the inner-fastest pair order, so at one zone they read one or a
// few contiguous runs of psi/rhs.  Every lane of a wave works on the same zone
// at the same time: the zone cursor, every branch and the zone part of every
// address are wave-uniform (scalar registers, soffset operands), and the
// lanes differ only by a per-pair base offset.
//
// Work split.  Wave w of a workgroup of KT waves owns k-layer kp*KT + w of
// k-patch kp and walks that layer's whole (j, i) plane: rows in j, zones in i.
// The x-face stays in a register, the j-face of each zone of the row in a
// register array indexed by i (the same lane meets zone (i, j+1) one row
// later), and the k-face goes to wave w+1 through LDS (double-buffered, one
// barrier per zone step); wave w runs w steps behind wave 0.  Faces never
// leave the chip inside a k-patch; between k-patches the top layer's faces go
// through the subdomain's own k-plane chunk.  A unit strings R pair groups
// (64 pairs each) back to back.
//
// Memory pipeline.  rhs, sigma_t and wave 0's incoming k-face of a step are
// read RING steps ahead into a register ring through buffer resources;
// across the end of a row the prefetch continues into the next row of the
// same pair group.  A lane whose pair lies past the subdomain's pairs reads
// and writes NOWHERE (zero, no access).
//
// Arithmetic per zone: the systolic kernel's (kr_sweep.hip), operation for
// operation, so psi is bit-identical across the two kernels and to the fma
// oracle:
//   num = fma(fz, c_k, fma(fy, b_j, fma(fx, a_i, rhs)))
//   psi = num / (((a_i + b_j) + c_k) + sigt)          (IEEE division)
//   f_out = 2*psi - f_in
#include "kr_launch.h"

#include <algorithm>

namespace kr {

namespace {

constexpr int WRING = 4;  // steps of prefetch (a multiple of it divides every instance's row length)

typedef __amdgpu_buffer_rsrc_t Rsrc;
constexpr uint32_t NOWHERE = 0x80000000u;  // resources are smaller than 2 GiB
constexpr int RSRC_WORD3 = 0x00020000;      // raw 32-bit dword data, no swizzle
constexpr int NT = 2;                       // cache policy: non-temporal

__device__ __forceinline__ Rsrc rsrc(const void *p, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)bytes, RSRC_WORD3);
}
// voffset: the lane's part (VGPR), soff: the wave-uniform zone part (SGPR)
template <int AUX>
__device__ __forceinline__ double ldw(Rsrc r, uint32_t voff, uint32_t soff) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, AUX));
}
template <int AUX>
__device__ __forceinline__ void stw(Rsrc r, uint32_t voff, uint32_t soff, double v) {
  typedef unsigned v2u __attribute__((ext_vector_type(2)));
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(v2u, v), r, voff, soff, AUX);
}

__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Wave-uniform position of one row: zone index of its first zone in sweep
// order and the face indices that stay fixed along the row.
struct RowPos {
  int32_t r, kp, jr;  // pair group, k-patch, row within the layer (sweep order)
  int32_t k, j;       // zone coordinates of the row
  bool valid;
};

__device__ __forceinline__ void row_coords(RowPos &rp, int kt, int KT, int nj, int nk, int jd, int kd) {
  const int kl = rp.kp * KT + kt;
  rp.k = kd > 0 ? kl : nk - 1 - kl;
  rp.j = jd > 0 ? rp.jr : nj - 1 - rp.jr;
}

__device__ __forceinline__ void row_next(RowPos &rp, int R, int NPK, int nj) {
  if (++rp.jr < nj) return;
  rp.jr = 0;
  if (++rp.kp < NPK) return;
  rp.kp = 0;
  ++rp.r;
  rp.valid = rp.r < R;
}

}  // namespace

template <int NI, bool POP, bool ADJ>
__global__ __launch_bounds__(1024) void k_sweep_waves(const DevSdom *__restrict__ sd,
                                                      const SweepPlanes *__restrict__ planes,
                                                      const SweepUnit *__restrict__ units, SweepShape sh,
                                                      double *__restrict__ pop) {
  extern __shared__ double lds[];
  const SweepUnit U = units[blockIdx.x];
  const DevSdom &S = sd[U.sdom];
  const SweepPlanes &PL = planes[U.sdom];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int kt = uni(tid >> 6);
  const int KT = sh.PK, NPK = sh.NPK;
  const int nd = S.nd, ng = S.ng, nj = S.nj, nk = S.nk;
  constexpr int ni = NI;  // the launcher picks the instance with NI == S.ni
  const int R = U.pg_count;
  const int jd = ADJ ? -S.jd : S.jd, kd = ADJ ? -S.kd : S.kd;
  const bool iup = (S.id > 0) != ADJ;
  const int L = ni + nj + nk + 1;  // coefficient record per direction: A[ni], B[nj], C[nk], w
  const int nt = blockDim.x;

  double *fzb = lds;               // [2][KT][64] k-face hand-off
  double *co = lds + 2 * nt;       // [nd][L]
  {
    const gmem<double> *src = gp(S.coef[0]);  // A[nd][ni], B[nd][nj], C[nd][nk], contiguous
    for (int e = tid; e < nd * (ni + nj + nk); e += nt) {
      int dst;
      if (e < nd * ni) {
        // A in sweep order: the zone a step reads sits at a fixed offset
        const int d = e / ni, x = e - d * ni;
        dst = d * L + (iup ? x : ni - 1 - x);
      } else if (e < nd * (ni + nj)) {
        const int d = (e - nd * ni) / nj;
        dst = d * L + ni + (e - nd * ni - d * nj);
      } else {
        const int d = (e - nd * (ni + nj)) / nk;
        dst = d * L + ni + nj + (e - nd * (ni + nj) - d * nk);
      }
      co[dst] = src[e];
    }
    if (POP)
      for (int d = tid; d < nd; d += nt) co[d * L + L - 1] = gp(S.w)[d];
  }
  __syncthreads();

  // resources based at the unit's first pair (offsets grow with the pair
  // index in every nesting; the host keeps each unit's span below 2 GiB)
  const int64_t npairs = (int64_t)nd * ng;
  const int64_t p0 = (int64_t)U.pg_begin * 64;
  const int d0 = sh.inner_is_g ? (int)(p0 / ng) : (int)(p0 % nd);
  const int g0 = sh.inner_is_g ? (int)(p0 % ng) : (int)(p0 / nd);
  const int64_t nz = (int64_t)ni * nj * nk;
  const int64_t nf[3] = {(int64_t)nj * nk, (int64_t)ni * nk, (int64_t)ni * nj};
  auto span = [](int64_t elems) { return (uint32_t)std::min<int64_t>(elems * 8, 0x7FFFFFF8LL); };
  const int64_t bf = d0 * S.fa + g0 * S.fg;
  const double *srhs = ADJ ? S.psi : S.rhs;
  double *spsi = ADJ ? S.rhs : S.psi;
  const Rsrc Rrhs = rsrc(srhs + bf, span(npairs * nz - bf));
  const Rsrc Rpsi = rsrc(spsi + bf, span(npairs * nz - bf));
  const Rsrc Rsig = rsrc(S.sig, span((int64_t)ng * nz));
  Rsrc Rin[3], Rout[3], Rown[3];
#pragma unroll
  for (int x = 0; x < 3; ++x) {
    const int64_t bp = d0 * S.pa[x] + g0 * S.pg[x];
    const uint32_t sp = span(npairs * nf[x] - bp);
    Rin[x] = rsrc(PL.in[x] ? PL.in[x] + bp : nullptr, PL.in[x] ? sp : 0u);  // vacuum: reads zeros
    Rout[x] = rsrc(PL.out[x] + bp, sp);
    Rown[x] = rsrc(S.plane[x] + bp, sp);
  }
  const bool has_in2 = PL.in[2] != nullptr;

  // per-lane parts of the offsets for pair group r of the unit
  uint32_t vf = NOWHERE, vs = NOWHERE, vp[3] = {NOWHERE, NOWHERE, NOWHERE};
  int cd = 0;  // the lane's direction: its coefficient record is co + cd*L
  auto lane_pair = [&](int r) {
    int64_t p = p0 + (int64_t)r * 64 + lane;
    const bool live = p < npairs;
    if (!live) p = npairs - 1;
    const int d = sh.inner_is_g ? (int)(p / ng) : (int)(p % nd);
    const int g = sh.inner_is_g ? (int)(p % ng) : (int)(p / nd);
    cd = d;
    const int64_t dd = d - d0, dg = g - g0;
    vf = live ? (uint32_t)(8 * (dd * S.fa + dg * S.fg)) : NOWHERE;
    vs = live ? (uint32_t)(8 * ((int64_t)g * S.sg)) : NOWHERE;
#pragma unroll
    for (int x = 0; x < 3; ++x) vp[x] = live ? (uint32_t)(8 * (dd * S.pa[x] + dg * S.pg[x])) : NOWHERE;
  };

  // wave-uniform byte offsets of a row's first zone in sweep order, and the
  // per-zone steps along the row (all 32-bit: every span is below 2 GiB)
  const int zs = iup ? 1 : -1;
  const int32_t dz8 = zs * 8 * (int32_t)S.fz, ds8 = zs * 8 * (int32_t)S.sz, dv8 = zs * 8;
  const int32_t dy8 = zs * 8 * (int32_t)S.pf[1], df8 = zs * 8 * (int32_t)S.pf[2];
  const int zi0 = iup ? 0 : ni - 1;
  struct RowOff {
    uint32_t z8, s8, v8;  // psi/rhs, sigma_t, volume at the row's first zone
    uint32_t x8, y8, f8;  // i-plane face of the row; j-plane and k-plane faces of its first zone
  };
  auto row_off = [&](const RowPos &rp) {
    const uint32_t z0 = (uint32_t)((rp.k * nj + rp.j) * ni + zi0);
    RowOff o;
    o.z8 = z0 * (8u * (uint32_t)S.fz);
    o.s8 = z0 * (8u * (uint32_t)S.sz);
    o.v8 = 8u * z0;
    o.x8 = (uint32_t)(rp.k * nj + rp.j) * (8u * (uint32_t)S.pf[0]);
    o.y8 = (uint32_t)(rp.k * ni + zi0) * (8u * (uint32_t)S.pf[1]);
    o.f8 = (uint32_t)(rp.j * ni + zi0) * (8u * (uint32_t)S.pf[2]);
    return o;
  };

  double rq[WRING], rs[WRING], rz[WRING];
#pragma unroll
  for (int s = 0; s < WRING; ++s) rq[s] = rs[s] = rz[s] = 0.0;
  // running prefetch position: offsets of the zone WRING steps ahead, stepped
  // by additions (one scalar add per stream and step)
  uint32_t pz8 = 0, ps8 = 0, pf8 = 0;
  bool pkf = false;  // wave 0 reads its k-face from the own k-plane (incoming face, or the previous patch's top)
  auto fetch = [&](int slot) {
    rq[slot] = ldw<NT>(Rrhs, vf, pz8);
    rs[slot] = ldw<0>(Rsig, vs, ps8);
    rz[slot] = pkf ? ldw<0>(Rown[2], vp[2], pf8) : 0.0;
    pz8 += (uint32_t)dz8;
    ps8 += (uint32_t)ds8;
    pf8 += (uint32_t)df8;
  };
  auto fetch_at = [&](const RowOff &o, bool kf) {
    pz8 = o.z8;
    ps8 = o.s8;
    pf8 = o.f8;
    pkf = kf;
  };

  const int total_rows = R * NPK * nj;
  double fy[NI];
  double pacc = 0.0;
  double wdir = 0.0;
  const bool vol_uniform = POP && cp(S.vol_info)[0] != 0.0;  // k_vol_info, run before every fused sweep
  const double vol0 = POP ? cp(S.vol_info)[1] : 0.0;

  // idle steps before this wave's first zone
  for (int s = 0; s < kt; ++s) __syncthreads();

  // a pair group's start: lane offsets, ring, the first row's x-face and j-faces
  auto group_start = [&](const RowPos &rp, double &fxn) {
    lane_pair(rp.r);
    if (POP) wdir = co[cd * L + L - 1];
    const RowOff o = row_off(rp);
    fetch_at(o, kt == 0 && (rp.kp > 0 || has_in2));
#pragma unroll
    for (int s = 0; s < WRING; ++s) fetch(s);
    fxn = ldw<0>(Rin[0], vp[0], o.x8);
    uint32_t y8 = o.y8;
#pragma unroll
    for (int i = 0; i < NI; ++i) {
      fy[i] = ldw<0>(Rin[1], vp[1], y8);
      y8 += (uint32_t)dy8;
    }
  };

  RowPos cur{0, 0, 0, 0, 0, true};
  row_coords(cur, kt, KT, nj, nk, jd, kd);
  double fxn = 0.0;  // the next row's incoming x-face
  group_start(cur, fxn);
  int par = kt & 1;  // LDS buffer this wave writes at its current step
  for (int row = 0; row < total_rows; ++row) {
    RowPos nx = cur;
    row_next(nx, R, NPK, nj);
    if (nx.valid) row_coords(nx, kt, KT, nj, nk, jd, kd);
    const bool new_group = nx.valid && nx.r != cur.r;  // no prefetch across pair groups
    const bool pre = nx.valid && !new_group;            // the next row's loads go out during this row
    const RowOff oc = row_off(cur);
    const bool ylast = cur.jr == nj - 1;                // the layer's last row: its j-faces leave
    const bool ktop = kt == KT - 1;
    const Rsrc Rk = cur.kp == NPK - 1 ? Rout[2] : Rown[2];  // the top layer's k-faces: downwind, or the next patch
    const double *rec = co + cd * L;
    const double bj = rec[ni + cur.j];
    const double ck = rec[ni + nj + cur.k];
    double fx = fxn;
    uint32_t ny8 = 0;
    if (pre) {
      const RowOff on = row_off(nx);
      fxn = ldw<0>(Rin[0], vp[0], on.x8);
      ny8 = on.y8;
    }
    uint32_t z8 = oc.z8, y8 = oc.y8, f8 = oc.f8, v8 = oc.v8;
    const int src = (kt > 0 ? kt - 1 : 0) * 64 + lane, dst = kt * 64 + lane;
#pragma unroll
    for (int i = 0; i < NI; ++i) {
      const int slot = i % WRING;
      const double a = rec[i];
      const double fzl = fzb[(par ^ 1) * nt + src];
      const double fz = kt == 0 ? rz[slot] : fzl;
      const double wx = ADJ ? 2.0 : a, wy = ADJ ? 2.0 : bj, wz = ADJ ? 2.0 : ck;
      const double num = __fma_rn(fz, wz, __fma_rn(fy[i], wy, __fma_rn(fx, wx, rq[slot])));
      const double den = ((a + bj) + ck) + rs[slot];
      const double p = num / den;
      const double p2 = 2.0 * p;
      fx = (ADJ ? a * p : p2) - fx;
      const double fyo = (ADJ ? bj * p : p2) - fy[i];
      const double fzo = (ADJ ? ck * p : p2) - fz;
      fy[i] = fyo;
      stw<NT>(Rpsi, vf, z8, p);
      if (ylast) {
        // the register takes the next layer's incoming face, needed NI - i steps from now
        stw<0>(Rout[1], vp[1], y8, fyo);
        if (pre) fy[i] = ldw<0>(Rin[1], vp[1], ny8);
        ny8 += (uint32_t)dy8;
      }
      if (ktop)
        stw<0>(Rk, vp[2], f8, fzo);
      else
        fzb[par * nt + dst] = fzo;
      if (POP) {
        // zone volume: one value for the whole subdomain (the generator's
        // meshes: no load), else the zone's own
        double vz = vol0;
        if (!vol_uniform) vz = cp(S.volume)[v8 >> 3];
        pacc += (wdir * p) * vz;
        // keep the sum in this step: left alone the compiler defers the
        // whole row's terms to the row's end and spills the row's psi values
        asm volatile("" : "+v"(pacc));
      }
      // refill this slot with the step WRING ahead: this row, then the next one
      if (i + WRING == NI) {
        if (pre) fetch_at(row_off(nx), kt == 0 && (nx.kp > 0 || has_in2));
      }
      if (i + WRING < NI || pre) fetch(slot);
      z8 += (uint32_t)dz8;
      y8 += (uint32_t)dy8;
      f8 += (uint32_t)df8;
      v8 += (uint32_t)dv8;
      par ^= 1;
      __syncthreads();
    }
    stw<0>(Rout[0], vp[0], oc.x8, fx);
    if (new_group) group_start(nx, fxn);
    cur = nx;
  }
  // idle steps after this wave's last zone
  for (int s = kt + 1; s < KT; ++s) __syncthreads();

  if (POP) {
    // fixed-shape tree over the block: the partial does not depend on timing
    lds[tid] = pacc;  // the face buffers are free after the last barrier
    __syncthreads();
    int half = 1;
    while (half * 2 < nt) half *= 2;
    for (; half > 0; half >>= 1) {
      if (tid < half && tid + half < nt) lds[tid] += lds[tid + half];
      __syncthreads();
    }
    if (tid == 0) pop[blockIdx.x] = lds[0];
  }
}

bool waves_shape(SweepShape &sh, int32_t nd, int32_t ng, int32_t ni, int32_t nj, int32_t nk, size_t lds_limit,
                 int kt_max) {
  const int64_t npairs = (int64_t)nd * ng;
  if (npairs % 64 != 0 || (ni != 8 && ni != 16 && ni != 32)) return false;  // the kernel's instances
  // one pair group's span of psi, of a plane and sigma_t in 32-bit byte offsets (host-checked per unit too)
  const int64_t nz = (int64_t)ni * nj * nk;
  if (npairs * nz * 8 >= (1LL << 31) || (int64_t)ng * nz * 8 >= (1LL << 31)) return false;
  for (int KT = std::min(kt_max, std::min(16, nk)); KT >= 1; --KT) {
    if (nk % KT != 0) continue;
    // between k-patches the top layer's faces go through the own plane and
    // are read back WRING steps before use, a whole layer later
    if (nk / KT > 1 && (int64_t)nj * ni < WRING + KT) continue;
    const size_t lds = (2 * (size_t)64 * KT + (size_t)nd * (ni + nj + nk + 1)) * sizeof(double);
    if (lds > lds_limit) continue;
    sh.kind = 3;
    sh.Q = 64;
    sh.PJ = 1;
    sh.PK = KT;
    sh.NPJ = 1;
    sh.NPK = nk / KT;
    return true;
  }
  return false;
}

int launch_sweep_waves(hipStream_t st, const DevSdom *sd, const SweepPlanes *planes, const SweepUnit *units,
                       int32_t num_units, const SweepShape &shape, double *pop, bool adjoint) {
  if (num_units <= 0) return KR_OK;
  const int nt = 64 * shape.PK;
  const size_t lds = (2 * (size_t)nt + (size_t)shape.nd * (shape.ni + shape.nj + shape.nk + 1)) * sizeof(double);
  const size_t lds_alloc = std::max(lds, (size_t)nt * sizeof(double));  // the population tree reuses it
#define KR_WAVES_LAUNCH(NIM)                                                                                     \
  do {                                                                                                           \
    if (lds_alloc > 65536) {                                                                                     \
      int rc;                                                                                                    \
      if ((rc = allow_lds((const void *)k_sweep_waves<NIM, false, true>, lds_alloc)) ||                          \
          (rc = allow_lds((const void *)k_sweep_waves<NIM, true, false>, lds_alloc)) ||                          \
          (rc = allow_lds((const void *)k_sweep_waves<NIM, false, false>, lds_alloc)))                           \
        return rc;                                                                                               \
    }                                                                                                            \
    if (adjoint)                                                                                                 \
      hipLaunchKernelGGL((k_sweep_waves<NIM, false, true>), dim3(num_units), dim3(nt), lds_alloc, st, sd, planes, \
                         units, shape, pop);                                                                     \
    else if (pop)                                                                                                \
      hipLaunchKernelGGL((k_sweep_waves<NIM, true, false>), dim3(num_units), dim3(nt), lds_alloc, st, sd, planes, \
                         units, shape, pop);                                                                     \
    else                                                                                                         \
      hipLaunchKernelGGL((k_sweep_waves<NIM, false, false>), dim3(num_units), dim3(nt), lds_alloc, st, sd,      \
                         planes, units, shape, pop);                                                             \
  } while (0)
  if (shape.ni == 8)
    KR_WAVES_LAUNCH(8);
  else if (shape.ni == 16)
    KR_WAVES_LAUNCH(16);
  else if (shape.ni == 32)
    KR_WAVES_LAUNCH(32);
  else {
    set_error("wave sweep: no instance for this row length");
    return KR_ERR_UNSUPPORTED;
  }
#undef KR_WAVES_LAUNCH
  KR_HIP_CHECK(hipGetLastError());
  return KR_OK;
}

}  // namespace kr
