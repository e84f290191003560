// Moment-space kernels of the source iteration on gfx950:
// LTimes, LPlusTimes, scattering, source, population, kConst/kCopy.
//
// All are HBM-streaming FP64 kernels.  One thread owns one (group, zone)
// column and keeps its moments in registers; the direction loop reads the
// quadrature tables through the scalar cache (they are wave-uniform).
// Accumulations follow the reference order per output element (directions
// ascending over subdomains ascending for LTimes, moments ascending for
// LPlusTimes) with fused multiply-adds.
#include "kr_launch.h"

#include <algorithm>

namespace kr {

// Blocks for n items at per_block per block, capped (grid-stride kernels).
static inline int grid_for(int64_t n, int per_block = 256, int max_blocks = 8192) {
  int64_t b = (n + per_block - 1) / per_block;
  if (b > max_blocks) b = max_blocks;
  if (b < 1) b = 1;
  return (int)b;
}

// Map a flat (group, zone) index to (g, z) so that consecutive threads are
// consecutive in memory for the field's stride pattern.
__device__ inline void gz_of(int64_t e, int32_t ng, int32_t nz, bool zfast, int32_t &g, int32_t &z) {
  if (zfast) {
    z = (int32_t)(e % nz);
    g = (int32_t)(e / nz);
  } else {
    g = (int32_t)(e % ng);
    z = (int32_t)(e / ng);
  }
}

// ---------------------------------------------------------------------------
// LTimes: src/Kripke/Kernel/LTimes.cpp:54-65 (body), :79-105 (sdom loop).
// A thread owns one (g,z) column and folds its psi into M accumulators in
// batches of DB directions, direction by direction, with ell read through
// the scalar cache (SGPR operands of the FMAs).  The loads of the next batch
// are issued before the FMAs of the current one (a two-batch software
// pipeline), so DB to 2 DB loads per thread are always in flight.
// ---------------------------------------------------------------------------
// ADJ: the same contraction as LPlusTimes' transpose: phi_out(nm) (+)= sum_d
// ell_plus(d,nm) rhs(d) (the adjoint of LPlusTimes maps rhs to phi_out).
template <int M, int DB, bool ACC, bool ADJ>
__global__ __launch_bounds__(256) void k_ltimes(const DevSdom *__restrict__ sd, const DevPR *__restrict__ prs,
                                                const int32_t *__restrict__ pr_sdoms) {
  const DevPR &P = prs[blockIdx.y];
  const int64_t n = (int64_t)P.ng * P.nz;
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const DevSdom &S0 = sd[pr_sdoms[P.sdom_begin]];
  const bool zfast = S0.fz <= S0.fg;
  const int nd = S0.nd;
  const int64_t fa = S0.fa;
  int32_t g, z;
  gz_of(e, P.ng, P.nz, zfast, g, z);
  const int64_t po = g * P.mg + z * P.mz;
  gmem<double> *phi = gp(ADJ ? P.phi_out : P.phi);
  double acc[M];
#pragma unroll
  for (int nm = 0; nm < M; ++nm) acc[nm] = ACC ? phi[po + nm * P.ma] : 0.0;
  // software pipeline over the batches (subdomain k, directions d0..d0+DB):
  // the next batch's loads are in flight while this batch's FMAs run
  const int nb = nd / DB, total = P.sdom_count * nb;  // DB divides nd
  auto batch_ptr = [&](int t, const cmem<double> *&ell) {
    const DevSdom &S = sd[pr_sdoms[P.sdom_begin + t / nb]];
    ell = cp(ADJ ? S.ell_plus : S.ell) + (t % nb) * DB * M;
    return gp(ADJ ? S.rhs : S.psi) + g * S0.fg + z * S0.fz + (int64_t)(t % nb) * DB * fa;
  };
  double cur[DB], nxt[DB];
  const cmem<double> *ell_cur, *ell_nxt = nullptr;
  {
    const gmem<double> *src = batch_ptr(0, ell_cur);
#pragma unroll
    for (int b = 0; b < DB; ++b) cur[b] = __builtin_nontemporal_load(src + b * fa);
  }
  for (int t = 0; t < total; ++t) {
    if (t + 1 < total) {
      const gmem<double> *src = batch_ptr(t + 1, ell_nxt);
#pragma unroll
      for (int b = 0; b < DB; ++b) nxt[b] = __builtin_nontemporal_load(src + b * fa);
    }
#pragma unroll
    for (int b = 0; b < DB; ++b) {
#pragma unroll
      for (int nm = 0; nm < M; ++nm) acc[nm] = __fma_rn(ell_cur[b * M + nm], cur[b], acc[nm]);
    }
#pragma unroll
    for (int b = 0; b < DB; ++b) cur[b] = nxt[b];
    ell_cur = ell_nxt;
  }
  // moment planes (ma > 1: the zone- and group-fastest nestings) take
  // streaming (non-temporal) stores: phi is not re-read by this kernel, and
  // keeping its M planes out of the caches leaves them to the psi stream.
  // Moments adjacent in memory (ma == 1) keep ordinary stores, which the
  // caches merge into whole lines.
  if (P.ma > 1) {
#pragma unroll
    for (int nm = 0; nm < M; ++nm) __builtin_nontemporal_store(acc[nm], phi + po + nm * P.ma);
  } else {
#pragma unroll
    for (int nm = 0; nm < M; ++nm) phi[po + nm] = acc[nm];
  }
}

// ---------------------------------------------------------------------------
// LPlusTimes: src/Kripke/Kernel/LPlusTimes.cpp:49-60, :68-92.
// A thread owns one (g,z) column with its M phi_out values in registers and
// produces DB directions at a time (DB independent FMA chains), with
// ell_plus read through the scalar cache.
// ---------------------------------------------------------------------------
// DB directions of one column, one FMA chain per direction (moments ascending).
template <int M, int DB, bool ACC>
__device__ __forceinline__ void lplus_batch(const double (&pho)[M], gmem<double> *rhs, int64_t fa,
                                            const cmem<double> *elp) {
  double acc[DB];
#pragma unroll
  for (int b = 0; b < DB; ++b) acc[b] = ACC ? rhs[b * fa] : 0.0;
#pragma unroll
  for (int b = 0; b < DB; ++b) {
#pragma unroll
    for (int nm = 0; nm < M; ++nm) acc[b] = __fma_rn(elp[b * M + nm], pho[nm], acc[b]);
  }
#pragma unroll
  for (int b = 0; b < DB; ++b) __builtin_nontemporal_store(acc[b], rhs + b * fa);
}

// ADJ: LTimes' transpose: psi(d) (+)= sum_nm ell(d,nm) phi(nm).
template <int M, int DB, bool ACC, bool ADJ>
__global__ __launch_bounds__(256) void k_lplustimes(const DevSdom *__restrict__ sd, const DevPR *__restrict__ prs,
                                                    const int32_t *__restrict__ pr_sdoms) {
  const DevPR &P = prs[blockIdx.y];
  const int64_t n = (int64_t)P.ng * P.nz;
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const DevSdom &S0 = sd[pr_sdoms[P.sdom_begin]];
  const bool zfast = S0.fz <= S0.fg;
  const int nd = S0.nd;
  const int64_t fa = S0.fa;
  int32_t g, z;
  gz_of(e, P.ng, P.nz, zfast, g, z);
  const int64_t po = g * P.mg + z * P.mz;
  const gmem<double> *pout = gp(ADJ ? P.phi : P.phi_out);
  double pho[M];
#pragma unroll
  for (int nm = 0; nm < M; ++nm) pho[nm] = pout[po + nm * P.ma];
  for (int k = 0; k < P.sdom_count; ++k) {
    const DevSdom &S = sd[pr_sdoms[P.sdom_begin + k]];
    gmem<double> *rhs = gp(ADJ ? S.psi : S.rhs) + g * S0.fg + z * S0.fz;
    const cmem<double> *elp = cp(ADJ ? S.ell : S.ell_plus);
    for (int d0 = 0; d0 < nd; d0 += DB) lplus_batch<M, DB, ACC>(pho, rhs + d0 * fa, fa, elp + d0 * M);  // DB divides nd
  }
}

// ---------------------------------------------------------------------------
// Direction-fastest nestings (GZD, ZGD): psi is [column][d] and phi is
// [column][nm] with the same column order (g*nz+z for GZD, z*ng+g for ZGD),
// so a thread-per-column mapping would touch one cache line per lane.  A
// block of CB columns moves its [CB][nd] and [CB][M] slabs through an LDS
// tile with coalesced loads and stores; each thread then runs its column from
// LDS with the same FMA order as the column kernels above.
// ---------------------------------------------------------------------------
// One wave per block: its barriers cost next to nothing.  On 64^3 zones x
// 128 groups x 128 directions (GZD) 64 columns took LPlusTimes from 8.7 to
// 8.0-8.2 ms and LTimes from 7.6-7.8 to 7.4 ms against 128; 256 took
// LPlusTimes to 10.9 ms.
constexpr int DF_CB = 64;  // columns per block

// Slab [n][w] (row-major, n <= DF_CB rows) <-> tile rows of stride ts.
// Element e = tid + i*DF_CB maps to (e / w, e % w), stepped without division.
struct SlabWalk {
  int col, k, dq, dr;
  __device__ SlabWalk(int tid, int w) : col(tid / w), k(tid % w), dq(DF_CB / w), dr(DF_CB % w) {}
  __device__ void next(int w) {
    col += dq;
    k += dr;
    if (k >= w) {
      k -= w;
      ++col;
    }
  }
};

template <bool NT>
__device__ __forceinline__ void slab_to_tile(double *tile, int ts, const gmem<double> *src, int total, int w) {
  constexpr int B = 8;  // loads in flight per thread before the LDS writes
  SlabWalk sw(threadIdx.x, w);
  for (int e0 = threadIdx.x; e0 < total; e0 += B * DF_CB) {
    double v[B];
#pragma unroll
    for (int b = 0; b < B; ++b) {
      const int e = e0 + b * DF_CB;
      v[b] = e < total ? (NT ? __builtin_nontemporal_load(src + e) : src[e]) : 0.0;
    }
#pragma unroll
    for (int b = 0; b < B; ++b) {
      if (e0 + b * DF_CB < total) tile[sw.col * ts + sw.k] = v[b];
      sw.next(w);
    }
  }
}

// One slab held in registers: element tid + b*DF_CB in v[b] (DF_PF*DF_CB >= total).
constexpr int DF_PF = 16;  // slab elements per thread of the register prefetch
__device__ __forceinline__ void slab_to_regs(double (&v)[DF_PF], const gmem<double> *src, int total) {
#pragma unroll
  for (int b = 0; b < DF_PF; ++b) {
    const int e = threadIdx.x + b * DF_CB;
    v[b] = e < total ? __builtin_nontemporal_load(src + e) : 0.0;
  }
}
__device__ __forceinline__ void regs_to_tile(double *tile, int ts, const double (&v)[DF_PF], int total, int w) {
  SlabWalk sw(threadIdx.x, w);
#pragma unroll
  for (int b = 0; b < DF_PF; ++b) {
    if (threadIdx.x + b * DF_CB < total) tile[sw.col * ts + sw.k] = v[b];
    sw.next(w);
  }
}

__device__ __forceinline__ void tile_to_slab(gmem<double> *dst, const double *tile, int ts, int total, int w) {
  SlabWalk sw(threadIdx.x, w);
  for (int e = threadIdx.x; e < total; e += DF_CB) {
    __builtin_nontemporal_store(tile[sw.col * ts + sw.k], dst + e);
    sw.next(w);
  }
}

template <int M, bool ACC, bool ADJ>
__global__ __launch_bounds__(DF_CB) void k_ltimes_dfast(const DevSdom *__restrict__ sd, const DevPR *__restrict__ prs,
                                                        const int32_t *__restrict__ pr_sdoms, int ts) {
  extern __shared__ double tile[];
  const DevPR &P = prs[blockIdx.y];
  const int64_t C = (int64_t)P.ng * P.nz;
  const int64_t c0 = (int64_t)blockIdx.x * DF_CB;
  if (c0 >= C) return;
  const int cn = (int)min<int64_t>(DF_CB, C - c0);
  const int tid = threadIdx.x;
  const int nd = sd[pr_sdoms[P.sdom_begin]].nd;
  gmem<double> *phi = gp(ADJ ? P.phi_out : P.phi) + c0 * M;
  double acc[M];
  if (ACC) {
    slab_to_tile<false>(tile, ts, phi, cn * M, M);
    __syncthreads();
#pragma unroll
    for (int nm = 0; nm < M; ++nm) acc[nm] = tid < cn ? tile[tid * ts + nm] : 0.0;
    __syncthreads();
  } else {
#pragma unroll
    for (int nm = 0; nm < M; ++nm) acc[nm] = 0.0;
  }
  auto slab = [&](int k) {
    const DevSdom &S = sd[pr_sdoms[P.sdom_begin + k]];
    return gp(ADJ ? S.rhs : S.psi) + c0 * nd;
  };
  auto fold = [&](int k) {
    const DevSdom &S = sd[pr_sdoms[P.sdom_begin + k]];
    const cmem<double> *ell = cp(ADJ ? S.ell_plus : S.ell);
    if (tid < cn) {
      for (int d = 0; d < nd; ++d) {
        const double p = tile[tid * ts + d];
#pragma unroll
        for (int nm = 0; nm < M; ++nm) acc[nm] = __fma_rn(ell[d * M + nm], p, acc[nm]);
      }
    }
  };
  if (nd <= DF_PF) {
    // the next subdomain's slab is loaded into registers while this one's
    // FMAs run from the tile, so the block's loads stay in flight
    double v[DF_PF];
    slab_to_regs(v, slab(0), cn * nd);
    for (int k = 0; k < P.sdom_count; ++k) {
      regs_to_tile(tile, ts, v, cn * nd, nd);
      __syncthreads();
      if (k + 1 < P.sdom_count) slab_to_regs(v, slab(k + 1), cn * nd);
      fold(k);
      __syncthreads();
    }
  } else {
    for (int k = 0; k < P.sdom_count; ++k) {
      slab_to_tile<true>(tile, ts, slab(k), cn * nd, nd);
      __syncthreads();
      fold(k);
      __syncthreads();
    }
  }
#pragma unroll
  for (int nm = 0; nm < M; ++nm) tile[tid * ts + nm] = acc[nm];
  __syncthreads();
  tile_to_slab(phi, tile, ts, cn * M, M);
}

template <int M, bool ACC, bool ADJ>
__global__ __launch_bounds__(DF_CB) void k_lplustimes_dfast(const DevSdom *__restrict__ sd,
                                                            const DevPR *__restrict__ prs,
                                                            const int32_t *__restrict__ pr_sdoms, int ts) {
  extern __shared__ double tile[];
  const DevPR &P = prs[blockIdx.y];
  const int64_t C = (int64_t)P.ng * P.nz;
  const int64_t c0 = (int64_t)blockIdx.x * DF_CB;
  if (c0 >= C) return;
  const int cn = (int)min<int64_t>(DF_CB, C - c0);
  const int tid = threadIdx.x;
  const int nd = sd[pr_sdoms[P.sdom_begin]].nd;
  slab_to_tile<false>(tile, ts, gp(ADJ ? P.phi : P.phi_out) + c0 * M, cn * M, M);
  __syncthreads();
  double pho[M];
#pragma unroll
  for (int nm = 0; nm < M; ++nm) pho[nm] = tid < cn ? tile[tid * ts + nm] : 0.0;
  __syncthreads();
  for (int k = 0; k < P.sdom_count; ++k) {
    const DevSdom &S = sd[pr_sdoms[P.sdom_begin + k]];
    gmem<double> *rhs = gp(ADJ ? S.psi : S.rhs) + c0 * nd;
    if (ACC) {
      slab_to_tile<false>(tile, ts, rhs, cn * nd, nd);
      __syncthreads();
    }
    const cmem<double> *elp = cp(ADJ ? S.ell : S.ell_plus);
    if (tid < cn) {
      for (int d = 0; d < nd; ++d) {
        double acc = ACC ? tile[tid * ts + d] : 0.0;
#pragma unroll
        for (int nm = 0; nm < M; ++nm) acc = __fma_rn(elp[d * M + nm], pho[nm], acc);
        tile[tid * ts + d] = acc;
      }
    }
    __syncthreads();
    tile_to_slab(rhs, tile, ts, cn * nd, nd);
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Scattering: src/Kripke/Kernel/Scattering.cpp:73-99.
// sigs_z(n,g,gp) = sum over the zone's mixed elements of sigs*fraction, in
// mixed-element order, product then sum (no fusion), as the reference.
// ---------------------------------------------------------------------------
struct SigsStrides {
  int64_t mat, n, g, gp;
};

__device__ inline double sigs_zone(const double *__restrict__ sigs, const SigsStrides &ss, const DevR &R, int32_t z,
                                   int64_t n, int64_t gg, int64_t ggp) {
  const int64_t m0 = R.zone_to_mixelem[z];
  const int32_t nmix = R.zone_to_num[z];
  double s = 0.0;
  for (int32_t q = 0; q < nmix; ++q) {
    const int64_t mix = m0 + q;
    const int64_t mat = R.mix_mat[mix];
    const double t = sigs[mat * ss.mat + n * ss.n + gg * ss.g + ggp * ss.gp] * R.mix_frac[mix];
    s = s + t;
  }
  return s;
}

// Structure of data/sigs, recomputed on every scattering call so that edits
// of the field between calls are honoured: each block ORs "an entry with
// n > 0 or g != gp is nonzero" over its share into flags[blockIdx.x].
__global__ __launch_bounds__(256) void k_sigs_structure(const double *__restrict__ sigs, SigsStrides ss, int64_t nmat,
                                                        int64_t L1, int64_t G, int32_t *__restrict__ flags) {
  __shared__ int any;
  if (threadIdx.x == 0) any = 0;
  __syncthreads();
  const int64_t n = nmat * L1 * G * G;
  int mine = 0;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t gp = e % G, g = (e / G) % G, nl = (e / (G * G)) % L1, m = e / (G * G * L1);
    if ((nl > 0 || g != gp) && sigs[m * ss.mat + nl * ss.n + g * ss.g + gp * ss.gp] != 0.0) mine = 1;
  }
  if (mine) atomicOr(&any, 1);
  __syncthreads();
  if (threadIdx.x == 0) flags[blockIdx.x] = any;
}

__device__ inline bool sigs_coupled(const int32_t *__restrict__ flags) {
  int v = 0;
  for (int b = 0; b < SIGS_FLAG_BLOCKS; ++b) v |= flags[b];
  return v != 0;
}

// Structure-aware path for isotropic, group-diagonal sigs (the generated
// data, src/Kripke/Generate/Data.cpp:87-101): every skipped product is an
// exact zero, so the result equals the dense loop bit for bit.
// ADJ: the transpose, phi (+)= S^T phi_out; a diagonal S is its own transpose.
template <bool ADJ>
__global__ __launch_bounds__(256) void k_scatter_diag(const DevPR *__restrict__ prs, const DevR *__restrict__ rs,
                                                      const double *__restrict__ sigs, SigsStrides ss, int M,
                                                      int accumulate, const int32_t *__restrict__ flags) {
  if (sigs_coupled(flags)) return;
  const DevPR &P = prs[blockIdx.y];
  const DevR &R = rs[P.r];
  const int64_t n = (int64_t)P.ng * P.nz;
  const bool zfast = P.mz <= P.mg;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    int32_t g, z;
    gz_of(e, P.ng, P.nz, zfast, g, z);
    const int64_t gg = P.glo + g;
    const double s = sigs_zone(sigs, ss, R, z, 0, gg, gg);
    const int64_t po = g * P.mg + z * P.mz;
    double *dst = ADJ ? P.phi : P.phi_out;
    const double *src = ADJ ? P.phi_out : P.phi;
    const double base0 = accumulate ? dst[po] : 0.0;
    const double t = s * src[po];
    dst[po] = base0 + t;
    for (int nm = 1; nm < M; ++nm) {
      const double b = accumulate ? dst[po + nm * P.ma] : 0.0;
      dst[po + nm * P.ma] = b + 0.0;
    }
  }
}

// k_scatter_diag, overwrite form, for the direction-fastest nestings (GZD,
// ZGD), where a column's M moments are adjacent in phi_out ([column][nm]):
// a block fills DF_CB columns' moments into an LDS tile and writes the
// [DF_CB][M] slab out with coalesced stores (a thread-per-column store of M
// adjacent values touches a different line per lane).  Same values as
// k_scatter_diag.
template <bool ADJ>
__global__ __launch_bounds__(DF_CB) void k_scatter_diag_dfast(const DevPR *__restrict__ prs,
                                                              const DevR *__restrict__ rs,
                                                              const double *__restrict__ sigs, SigsStrides ss, int M,
                                                              const int32_t *__restrict__ flags) {
  extern __shared__ double tile[];
  if (sigs_coupled(flags)) return;
  const DevPR &P = prs[blockIdx.y];
  const DevR &R = rs[P.r];
  const int64_t C = (int64_t)P.ng * P.nz;
  const int64_t c0 = (int64_t)blockIdx.x * DF_CB;
  if (c0 >= C) return;
  const int cn = (int)min<int64_t>(DF_CB, C - c0);
  const int ts = M | 1;
  const bool zfast = P.mz <= P.mg;
  double *dst = ADJ ? P.phi : P.phi_out;
  const double *src = ADJ ? P.phi_out : P.phi;
  const int tid = threadIdx.x;
  if (tid < cn) {
    int32_t g, z;
    gz_of(c0 + tid, P.ng, P.nz, zfast, g, z);
    const int64_t po = g * P.mg + z * P.mz;
    const int64_t gg = P.glo + g;
    const double t = sigs_zone(sigs, ss, R, z, 0, gg, gg) * src[po];
    tile[tid * ts] = 0.0 + t;
    for (int nm = 1; nm < M; ++nm) tile[tid * ts + nm] = 0.0 + 0.0;
  }
  __syncthreads();
  tile_to_slab(gp(dst) + c0 * M, tile, ts, cn * M, M);
}

// General sigs (group coupling, anisotropic moments).  The zone's matrix
// sigs_z(n,g,gp) = sum over its mixed elements of frac * sigs(mat,n,g,gp) is
// linear in the materials, so
//     phi_out(nm,g,z) = sum_mat W(z,mat) sum_gp sigs(mat,n,g,gp) phi(nm,gp,z)
// with W(z,mat) the zone's fraction of each material.  Per material this is
// a dense (groups x groups) x (groups x zones) product, mapped so that a
// wave's loads follow the nesting's fastest axis:
//   zone fastest (DGZ, GDZ)    k_scatter_dense: lane = zone, one moment per
//                              block, SG_OC output groups per lane;
//   moment fastest (GZD, ZGD)  k_scatter_dense_mfast: lane = (zone, moment)
//                              over the moments of one Legendre order, so
//                              the coefficient row is still wave-uniform;
//   group fastest (DZG, ZDG)   k_scatter_dense_gfast: lane = output group,
//                              GF_ZT zones per lane, phi of a zone read
//                              through the scalar cache.
// In the first two the coefficient rows come through the scalar cache (one
// load feeds SG_OC FMAs per loaded phi) and a wave skips every material none
// of its zones contains.  The association differs from the reference's
// per-zone matrix (rounding level).  k_sigs_relayout first writes the
// coefficients as T[mat][n][h][o] with the output group o fastest: forward
// o = g, h = gp; transpose o = gp, h = g.  T carries SG_PAD zeros at its end:
// a partial output tile reads up to SG_OC - 1 entries past its row.
constexpr int SG_OC = 16;
constexpr int SG_PAD = 64;
constexpr int GF_ZT = 16;  // zones per lane, group-fastest kernel
constexpr int GF_HB = 8;   // source groups per scalar load, group-fastest kernel
constexpr int MF_ZS = 252; // zones per supertile, moment-fastest kernel
// Dense launches cap their grids (the kernels loop over the rest): with an
// uncoupled sigs they return at once, and the launch then costs few blocks.
// The moment-fastest kernel runs best with fewer, longer blocks.
constexpr int64_t DENSE_MAX_BLOCKS = 16384;
constexpr int64_t DENSE_MAX_BLOCKS_MFAST = 4096;

static int64_t tab_doubles(int64_t L1, int64_t M) { return (L1 + 1 + M + 1) / 2 + 1; }

__global__ __launch_bounds__(256) void k_sigs_relayout(const double *__restrict__ sigs, SigsStrides ss, int64_t nmat,
                                                       int64_t L1, int64_t G, double *__restrict__ T, int adj,
                                                       const int32_t *__restrict__ flags) {
  if (!sigs_coupled(flags)) return;
  const int64_t n = nmat * L1 * G * G;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n + SG_PAD;
       e += (int64_t)gridDim.x * blockDim.x) {
    if (e >= n) {
      T[e] = 0.0;
      continue;
    }
    const int64_t o = e % G, h = (e / G) % G, nl = (e / (G * G)) % L1, m = e / (G * G * L1);
    const int64_t g = adj ? h : o, gp = adj ? o : h;
    T[e] = sigs[m * ss.mat + nl * ss.n + g * ss.g + gp * ss.gp];
  }
}

// Moments grouped by Legendre order from moment_to_legendre: tab[n] is the
// first slot of order n, tab[L1] the count, tab[L1 + 1 + slot] the moment.
// Moments whose order lies outside [0, L1) take no part.
__global__ void k_moment_orders(const int64_t *__restrict__ m2l, int M, int L1, int32_t *__restrict__ tab,
                                const int32_t *__restrict__ flags) {
  if (!sigs_coupled(flags) || threadIdx.x != 0 || blockIdx.x != 0) return;
  int k = 0;
  for (int n = 0; n < L1; ++n) {
    tab[n] = k;
    for (int nm = 0; nm < M; ++nm)
      if (m2l[nm] == n) tab[L1 + 1 + k++] = nm;
  }
  tab[L1] = k;
}

// W[r][mat][z] (row stride wstride): the fraction of material mat in zone z
// of zone set r, summed over the zone's mixed elements in their order.
__global__ __launch_bounds__(256) void k_mat_weights(const DevR *__restrict__ rs, int64_t nmat, int64_t wstride,
                                                     double *__restrict__ W, const int32_t *__restrict__ flags) {
  if (!sigs_coupled(flags)) return;
  const DevR &R = rs[blockIdx.y];
  double *Wr = W + (int64_t)blockIdx.y * nmat * wstride;
  for (int64_t z = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; z < R.nz; z += (int64_t)gridDim.x * blockDim.x) {
    for (int64_t m = 0; m < nmat; ++m) Wr[m * wstride + z] = 0.0;
    const int64_t m0 = R.zone_to_mixelem[z];
    const int32_t nmix = R.zone_to_num[z];
    for (int32_t q = 0; q < nmix; ++q) {
      const int64_t mat = R.mix_mat[m0 + q];
      if (mat >= 0 && mat < nmat) Wr[mat * wstride + z] += R.mix_frac[m0 + q];
    }
  }
}

// acc[o] = sum_mat W(z,mat) sum_h T[mat][nl][gh_lo + h][go_lo + o0 + o] * in[h*mg]
// for one lane (Wz = &W[r][0][z]); nl and the T offsets are wave-uniform.
__device__ __forceinline__ void dense_lane(double (&acc)[SG_OC], const double *__restrict__ Wz, int64_t wstride,
                                           bool live, const double *__restrict__ T, int64_t toff, int64_t nl,
                                           int64_t nmat, int64_t L1, int64_t G, const gmem<double> *in, int64_t mg,
                                           int nh) {
#pragma unroll
  for (int o = 0; o < SG_OC; ++o) acc[o] = 0.0;
  for (int64_t mat = 0; mat < nmat; ++mat) {
    const double wgt = live ? gp(Wz)[mat * wstride] : 0.0;
    if (__ballot(wgt != 0.0) == 0) continue;  // no zone of this wave has the material
    const cmem<double> *Tm = cp(T) + (mat * L1 + nl) * G * G + toff;
    int h = 0;
    for (; h + 4 <= nh; h += 4) {  // four loads in flight per step
      double p[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) p[u] = in[(int64_t)(h + u) * mg];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const double x = wgt * p[u];
        const cmem<double> *row = Tm + (int64_t)(h + u) * G;
#pragma unroll
        for (int o = 0; o < SG_OC; ++o) acc[o] = __fma_rn(row[o], x, acc[o]);
      }
    }
    for (; h < nh; ++h) {
      const double x = wgt * in[(int64_t)h * mg];
      const cmem<double> *row = Tm + (int64_t)h * G;
#pragma unroll
      for (int o = 0; o < SG_OC; ++o) acc[o] = __fma_rn(row[o], x, acc[o]);
    }
  }
}

__device__ __forceinline__ void dense_store(const double (&acc)[SG_OC], double *out, int64_t base, int o0, int ng,
                                            int64_t mg, int accumulate) {
#pragma unroll
  for (int o = 0; o < SG_OC; ++o) {
    if (o0 + o < ng) {
      const int64_t od = base + (int64_t)(o0 + o) * mg;
      out[od] = (accumulate ? out[od] : 0.0) + acc[o];
    }
  }
}

// Per (src, dst) chunk pair.  ADJ: the transpose for the same pair:
// phi_src(nm,gp,z) (+)= sum over the destination chunk's groups g of
// sigs_z(n,g,gp) * phi_out_dst(nm,g,z).
#define KR_DENSE_PAIR                                                           \
  const DevPR &Ps = prs[src];                                                   \
  const DevPR &Pd = prs[dst];                                                   \
  const double *Wr = W + (int64_t)Pd.r * nmat * wstride;                        \
  const DevPR &Po = ADJ ? Ps : Pd; /* chunk owning the outputs */               \
  const DevPR &Pi = ADJ ? Pd : Ps; /* chunk read */                             \
  const int64_t go_lo = ADJ ? Ps.glo : Pd.glo, gh_lo = ADJ ? Pd.glo : Ps.glo;   \
  double *out = ADJ ? Po.phi : Po.phi_out;                                      \
  const double *inb = ADJ ? Pi.phi_out : Pi.phi;

// Zone fastest: grid (zone blocks, moments x output tiles).
template <bool ADJ>
__global__ __launch_bounds__(256) void k_scatter_dense(const DevPR *__restrict__ prs, const double *__restrict__ W,
                                                       int64_t wstride, const double *__restrict__ T,
                                                       const int64_t *__restrict__ m2l,
                                                       int M, int64_t nmat, int64_t L1, int64_t G, int src, int dst,
                                                       int accumulate, const int32_t *__restrict__ flags) {
  if (!sigs_coupled(flags)) return;
  KR_DENSE_PAIR
  const int noc = (Po.ng + SG_OC - 1) / SG_OC;
  for (int yb = blockIdx.y; yb < M * noc; yb += gridDim.y) {  // the grid is capped in y
    const int nm = yb / noc;
    const int o0 = (yb - nm * noc) * SG_OC;
    const int64_t nl = m2l[nm];
    if (nl < 0 || nl >= L1) continue;
    const int z = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = z < Po.nz;
    const int zc = live ? z : 0;
    double acc[SG_OC];
    dense_lane(acc, Wr + zc, wstride, live, T, gh_lo * G + go_lo + o0, nl, nmat, L1, G,
               gp(inb) + nm * Pi.ma + zc * Pi.mz, Pi.mg, Pi.ng);
    if (live) dense_store(acc, out, nm * Po.ma + (int64_t)z * Po.mz, o0, Po.ng, Po.mg, accumulate);
  }
}

// Moment fastest: a wave covers zpw = 64 / mc zones x mc moments of one
// order (mc = min(count, 64); orders with more than 64 moments take several
// moment chunks), so its coefficient row is wave-uniform.  The zones are cut
// into supertiles of MF_ZS; one block takes one (supertile, order, output
// tile).  All (order, tile) blocks of a supertile read the same phi lines, so
// the 1-D grid is laid out XCD-aware: blocks b, b + 8, b + 16, ... run on one
// XCD, and consecutive blocks there take the nbt = orders x tiles blocks of
// one supertile, which then share that XCD's L2.
template <bool ADJ>
__global__ __launch_bounds__(256) void k_scatter_dense_mfast(const DevPR *__restrict__ prs,
                                                             const double *__restrict__ W, int64_t wstride,
                                                             const double *__restrict__ T,
                                                             const int32_t *__restrict__ tab, int64_t nmat,
                                                             int64_t L1, int64_t G, int src, int dst, int accumulate,
                                                             const int32_t *__restrict__ flags, int nocg,
                                                             int64_t nblk) {
  if (!sigs_coupled(flags)) return;
  KR_DENSE_PAIR
  const int64_t nbt = L1 * nocg;
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  // the grid is capped: a block takes work blocks bi, bi + gridDim.x, ...
  // (gridDim.x is a multiple of 8, so each keeps its XCD)
  for (int64_t bi = blockIdx.x; bi < nblk; bi += gridDim.x) {
    const int64_t xi = bi >> 3;
    const int64_t st = (xi / nbt) * 8 + (bi & 7);
    const int k = (int)(xi % nbt);
    const int nl = k / nocg;
    const int o0 = (k - nl * nocg) * SG_OC;
    const int64_t zs0 = st * MF_ZS;
    if (zs0 >= Po.nz || o0 >= Po.ng) continue;
    const int64_t zs1 = (zs0 + MF_ZS < Po.nz) ? zs0 + MF_ZS : Po.nz;
    const int start = tab[nl], cnt = tab[nl + 1] - start;
    if (cnt <= 0) continue;
    const int mc = cnt < 64 ? cnt : 64, nmc = (cnt + 63) / 64, zpw = 64 / mc;
    const int zl = lane / mc, ml = lane - zl * mc;
    const int64_t nwork = ((zs1 - zs0 + zpw - 1) / zpw) * nmc;
    for (int64_t w = wave; w < nwork; w += 4) {
      const int64_t zg = w / nmc;
      const int mk = (int)(w - zg * nmc) * 64 + ml;
      const int64_t z = zs0 + zg * zpw + zl;
      const bool live = zl < zpw && z < zs1 && mk < cnt;
      const int nm = tab[L1 + 1 + start + (live ? mk : 0)];
      const int64_t zc = live ? z : 0;
      double acc[SG_OC];
      dense_lane(acc, Wr + zc, wstride, live, T, gh_lo * G + go_lo + o0, nl, nmat, L1, G,
                 gp(inb) + nm * Pi.ma + zc * Pi.mz, Pi.mg, Pi.ng);
      if (live) dense_store(acc, out, nm * Po.ma + z * Po.mz, o0, Po.ng, Po.mg, accumulate);
    }
  }
}

// Group fastest: grid (grid-stride over zone tiles of GF_ZT, moments x
// 64-group output tiles); lane = output group.  Per material, a tile none of
// whose zones holds it is skipped; the GF_HB phi values of one zone come in
// one scalar load and each loaded coefficient feeds GF_ZT FMAs.  Per zone the
// material's sum is formed first and then weighted: sum_mat W * (sum_h T phi).
template <bool ADJ>
__global__ __launch_bounds__(256) void k_scatter_dense_gfast(const DevPR *__restrict__ prs,
                                                             const double *__restrict__ W, int64_t wstride,
                                                             const double *__restrict__ T,
                                                             const int64_t *__restrict__ m2l, int M, int64_t nmat,
                                                             int64_t L1, int64_t G, int src, int dst, int accumulate,
                                                             const int32_t *__restrict__ flags) {
  if (!sigs_coupled(flags)) return;
  KR_DENSE_PAIR
  const int noc = (Po.ng + 63) / 64;
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t ntile = ((int64_t)Po.nz + GF_ZT - 1) / GF_ZT;
  const int nh = Pi.ng;
  for (int yb = blockIdx.y; yb < M * noc; yb += gridDim.y) {  // the grid is capped in y
  const int nm = yb / noc;
  const int o0 = (yb - nm * noc) * 64;
  const int64_t nl = m2l[nm];
  if (nl < 0 || nl >= L1) continue;
  const int o = o0 + lane;
  const bool olive = o < Po.ng;
  for (int64_t t = (int64_t)blockIdx.x * 4 + wave; t < ntile; t += (int64_t)gridDim.x * 4) {
    const int64_t z0 = t * GF_ZT;
    const int nzt = (int)((Po.nz - z0) < GF_ZT ? (Po.nz - z0) : GF_ZT);
    double acc[GF_ZT];
#pragma unroll
    for (int zt = 0; zt < GF_ZT; ++zt) acc[zt] = 0.0;
    for (int64_t mat = 0; mat < nmat; ++mat) {
      double wz[GF_ZT];
      bool any = false;
      const cmem<double> *wrow = cp(Wr) + mat * wstride + z0;  // wstride >= zones + GF_ZT: a tile's row stays inside
#pragma unroll
      for (int zt = 0; zt < GF_ZT; ++zt) {
        wz[zt] = zt < nzt ? wrow[zt] : 0.0;
        any |= wz[zt] != 0.0;
      }
      if (!any) continue;  // zone-uniform: no zone of the tile has the material
      double inner[GF_ZT];
#pragma unroll
      for (int zt = 0; zt < GF_ZT; ++zt) inner[zt] = 0.0;
      const gmem<double> *Tc = gp(T) + (mat * L1 + nl) * G * G + gh_lo * G + go_lo + (olive ? o : o0);
      const double *in0 = inb + nm * Pi.ma + z0 * Pi.mz;
      int h = 0;
      for (; h + GF_HB <= nh; h += GF_HB) {
        double tc[GF_HB];
#pragma unroll
        for (int u = 0; u < GF_HB; ++u) tc[u] = Tc[(int64_t)(h + u) * G];
#pragma unroll
        for (int zt = 0; zt < GF_ZT; ++zt) {
          const int zr = zt < nzt ? zt : 0;
          const cmem<double> *x = cp(in0) + (int64_t)zr * Pi.mz + h;
#pragma unroll
          for (int u = 0; u < GF_HB; ++u) inner[zt] = __fma_rn(tc[u], x[u], inner[zt]);
        }
      }
      for (; h < nh; ++h) {
        const double tc = Tc[(int64_t)h * G];
#pragma unroll
        for (int zt = 0; zt < GF_ZT; ++zt) {
          const int zr = zt < nzt ? zt : 0;
          inner[zt] = __fma_rn(tc, cp(in0)[(int64_t)zr * Pi.mz + h], inner[zt]);
        }
      }
#pragma unroll
      for (int zt = 0; zt < GF_ZT; ++zt) acc[zt] = __fma_rn(wz[zt], inner[zt], acc[zt]);
    }
    if (olive) {
#pragma unroll
      for (int zt = 0; zt < GF_ZT; ++zt) {
        if (zt < nzt) {
          const int64_t od = nm * Po.ma + (int64_t)o * Po.mg + (z0 + zt) * Po.mz;
          out[od] = (accumulate ? out[od] : 0.0) + acc[zt];
        }
      }
    }
  }
  }
}
#undef KR_DENSE_PAIR

// Scattering with a per-zone, isotropic, group-diagonal cross section s(g,z)
// given directly (one array per PR chunk in sigt_zonal layout) instead of
// through materials: phi_out(0,g,z) (+)= s(g,z)*phi(0,g,z), higher moments
// (+)= 0.  The form the differentiable solver uses for per-zone sigma_s; for
// s equal to the mixed-element average of the generated sigs it matches
// k_scatter_diag.  ADJ: the transpose (phi from phi_out), the same map.
template <bool ADJ>
__global__ __launch_bounds__(256) void k_scatter_zonal(const DevPR *__restrict__ prs, const double *const *__restrict__ sz,
                                                       int64_t s_g, int64_t s_z, int M, int accumulate) {
  const DevPR &P = prs[blockIdx.y];
  const double *sc = sz[blockIdx.y];
  const int64_t n = (int64_t)P.ng * P.nz;
  const bool zfast = P.mz <= P.mg;
  double *dst = ADJ ? P.phi : P.phi_out;
  const double *src = ADJ ? P.phi_out : P.phi;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    int32_t g, z;
    gz_of(e, P.ng, P.nz, zfast, g, z);
    const int64_t po = g * P.mg + z * P.mz;
    const double t = sc[g * s_g + z * s_z] * src[po];
    dst[po] = (accumulate ? dst[po] : 0.0) + t;
    for (int nm = 1; nm < M; ++nm) dst[po + nm * P.ma] = (accumulate ? dst[po + nm * P.ma] : 0.0) + 0.0;
  }
}

// ---------------------------------------------------------------------------
// Parameter gradients of the differentiable source iteration (no reference
// counterpart).  With psi' = H(sigt)^-1 rhs, d psi'/d sigt(g,z) = -H^-1 diag
// (psi'), so dJ/d sigt(g,z) = -sum_{subdomains, d} rhsbar(d,g,z) psi'(d,g,z),
// where rhsbar = H^-T psibar is what the adjoint sweep leaves in the adjoint
// problem's rhs.  For the zonal scattering phi_out(0) = s phi(0):
// dJ/ds(g,z) = phi_outbar(0,g,z) phi(0,g,z).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_grad_sigt(const DevSdom *__restrict__ fsd, const DevSdom *__restrict__ asd,
                                                   const DevPR *__restrict__ prs, const int32_t *__restrict__ pr_sdoms,
                                                   double *const *__restrict__ grad, int64_t s_g, int64_t s_z,
                                                   int accumulate) {
  const DevPR &P = prs[blockIdx.y];
  const int64_t n = (int64_t)P.ng * P.nz;
  const DevSdom &S0 = fsd[pr_sdoms[P.sdom_begin]];
  const bool zfast = S0.fz <= S0.fg;
  double *gr = grad[blockIdx.y];
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    int32_t g, z;
    gz_of(e, P.ng, P.nz, zfast, g, z);
    const int64_t o = g * S0.fg + z * S0.fz;
    double acc = 0.0;
    for (int k = 0; k < P.sdom_count; ++k) {
      const int sdom = pr_sdoms[P.sdom_begin + k];
      const gmem<double> *pf = gp(fsd[sdom].psi) + o;
      const gmem<double> *ra = gp(asd[sdom].rhs) + o;
      for (int d = 0; d < S0.nd; ++d) acc = __fma_rn(ra[d * S0.fa], pf[d * S0.fa], acc);
    }
    const int64_t go = g * s_g + z * s_z;
    gr[go] = (accumulate ? gr[go] : 0.0) - acc;
  }
}

__global__ __launch_bounds__(256) void k_grad_sigs_zonal(const DevPR *__restrict__ fpr, const DevPR *__restrict__ apr,
                                                         double *const *__restrict__ grad, int64_t s_g, int64_t s_z,
                                                         int accumulate) {
  const DevPR &F = fpr[blockIdx.y];
  const DevPR &A = apr[blockIdx.y];
  const int64_t n = (int64_t)F.ng * F.nz;
  const bool zfast = F.mz <= F.mg;
  double *gr = grad[blockIdx.y];
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    int32_t g, z;
    gz_of(e, F.ng, F.nz, zfast, g, z);
    const int64_t po = g * F.mg + z * F.mz;
    const int64_t go = g * s_g + z * s_z;
    gr[go] = (accumulate ? gr[go] : 0.0) + A.phi_out[po] * F.phi[po];
  }
}

int launch_grad_sigt(const LaunchCtx &F, const DevSdom *adj_sd, double *const *grad_dev, int64_t s_g, int64_t s_z,
                     int accumulate) {
  dim3 grid(grid_for(F.max_gz), F.num_pr);
  hipLaunchKernelGGL(k_grad_sigt, grid, dim3(256), 0, F.stream, F.sd, adj_sd, F.pr, F.pr_sdoms, grad_dev, s_g, s_z,
                     accumulate);
  KR_HIP_CHECK(hipGetLastError());
  return KR_OK;
}

int launch_grad_sigs_zonal(const LaunchCtx &F, const DevPR *adj_pr, double *const *grad_dev, int64_t s_g, int64_t s_z,
                           int accumulate) {
  dim3 grid(grid_for(F.max_gz), F.num_pr);
  hipLaunchKernelGGL(k_grad_sigs_zonal, grid, dim3(256), 0, F.stream, F.pr, adj_pr, grad_dev, s_g, s_z, accumulate);
  KR_HIP_CHECK(hipGetLastError());
  return KR_OK;
}

// ---------------------------------------------------------------------------
// Source: src/Kripke/Kernel/Source.cpp:59-75.  A zone has at most one
// material-0 mixed element, so the updates never collide.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_source(const DevPR *__restrict__ prs, const DevR *__restrict__ rs,
                                                double strength) {
  const DevPR &P = prs[blockIdx.y];
  const DevR &R = rs[P.r];
  const int64_t n = (int64_t)P.ng * R.num_mix;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t mix = e % R.num_mix;
    const int32_t g = (int32_t)(e / R.num_mix);
    if (R.mix_mat[mix] == 0) {
      const int64_t z = R.mix_zone[mix];
      double *p = P.phi_out + g * P.mg + z * P.mz;
      *p = *p + strength * R.mix_frac[mix];
    }
  }
}

// ---------------------------------------------------------------------------
// Multi-problem forms for the batched launch: one launch covers the PR chunks
// of many problems, each chunk carrying its problem's scattering data.
// ---------------------------------------------------------------------------
// k_sigs_structure for every problem of a batch: block (b, p) classifies
// share b of problem p's sigs into its flags[b].
__global__ __launch_bounds__(256) void k_sigs_structure_multi(const DevScat *__restrict__ probs) {
  const DevScat &Q = probs[blockIdx.y];
  __shared__ int any;
  if (threadIdx.x == 0) any = 0;
  __syncthreads();
  const int64_t n = Q.nmat * Q.L1 * Q.G * Q.G;
  int mine = 0;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t gp_ = e % Q.G, g = (e / Q.G) % Q.G, nl = (e / (Q.G * Q.G)) % Q.L1, m = e / (Q.G * Q.G * Q.L1);
    if ((nl > 0 || g != gp_) && Q.sigs[m * Q.str[0] + nl * Q.str[1] + g * Q.str[2] + gp_ * Q.str[3]] != 0.0) mine = 1;
  }
  if (mine) atomicOr(&any, 1);
  __syncthreads();
  if (threadIdx.x == 0) Q.flags[blockIdx.x] = any;
}

// k_scatter_diag over the chunks of many problems (overwrite form); chunks
// of problems whose sigs is coupled return (their problem's own dense
// launches follow).
__global__ __launch_bounds__(256) void k_scatter_diag_multi(const DevPR *__restrict__ prs, const DevR *__restrict__ rs,
                                                            const DevScat *__restrict__ cs) {
  const DevScat &Q = cs[blockIdx.y];
  if (sigs_coupled(Q.flags)) return;
  const DevPR &P = prs[blockIdx.y];
  const DevR &R = rs[P.r];
  const SigsStrides ss{Q.str[0], Q.str[1], Q.str[2], Q.str[3]};
  const int64_t n = (int64_t)P.ng * P.nz;
  const bool zfast = P.mz <= P.mg;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    int32_t g, z;
    gz_of(e, P.ng, P.nz, zfast, g, z);
    const int64_t gg = P.glo + g;
    const double sz = sigs_zone(Q.sigs, ss, R, z, 0, gg, gg);
    const int64_t po = g * P.mg + z * P.mz;
    P.phi_out[po] = 0.0 + sz * P.phi[po];
    for (int nm = 1; nm < Q.M; ++nm) P.phi_out[po + nm * P.ma] = 0.0 + 0.0;
  }
}

// k_source over the chunks of many problems, each with its strength.
__global__ __launch_bounds__(256) void k_source_multi(const DevPR *__restrict__ prs, const DevR *__restrict__ rs,
                                                      const DevScat *__restrict__ cs) {
  const DevPR &P = prs[blockIdx.y];
  const DevR &R = rs[P.r];
  const double strength = cs[blockIdx.y].strength;
  const int64_t n = (int64_t)P.ng * R.num_mix;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t mix = e % R.num_mix;
    const int32_t g = (int32_t)(e / R.num_mix);
    if (R.mix_mat[mix] == 0) {
      const int64_t z = R.mix_zone[mix];
      double *q = P.phi_out + g * P.mg + z * P.mz;
      *q = *q + strength * R.mix_frac[mix];
    }
  }
}

// Per problem p (block p): the sum of its sweep-unit population partials,
// listed in its own unit order, with k_pop_final's pattern, so the value is
// the same as that problem's own fused population.
__global__ __launch_bounds__(256) void k_pop_sum_multi(const double *__restrict__ partial,
                                                       const int64_t *__restrict__ list,
                                                       const int64_t *__restrict__ begin, double *__restrict__ out,
                                                       int64_t out_stride) {
  __shared__ double red[256];
  const int64_t b0 = begin[blockIdx.x], n = begin[blockIdx.x + 1] - b0;
  double acc = 0.0;
  for (int64_t e = threadIdx.x; e < n; e += blockDim.x) acc += partial[list[b0 + e]];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[(int64_t)blockIdx.x * out_stride] = red[0];
}

// Per R chunk (block b): whether every zone volume equals the first (bitwise
// ==, so NaN never counts as uniform), and that volume.  The fused population
// of the row-scan sweep then takes the volume from this pair instead of
// loading it per zone (the reference's generator makes it one value,
// src/Kripke/Generate/Space.cpp:133-136).
__global__ __launch_bounds__(1024) void k_vol_info(const DevR *__restrict__ rs) {
  const DevR &R = rs[blockIdx.x];
  __shared__ int differs;
  if (threadIdx.x == 0) differs = 0;
  __syncthreads();
  const double v0 = R.volume[0];
  int mine = !(v0 == v0);
  for (int64_t z = threadIdx.x; z < R.nz && !mine; z += blockDim.x) mine = !(R.volume[z] == v0);
  if (mine) atomicOr(&differs, 1);
  __syncthreads();
  if (threadIdx.x == 0) {
    R.vol_info[0] = differs ? 0.0 : 1.0;
    R.vol_info[1] = v0;
  }
}

// ---------------------------------------------------------------------------
// Population: src/Kripke/Kernel/Population.cpp:49-63.  Term order per
// element is (w*psi)*volume; the sum is a fixed two-level tree, so the
// result does not depend on scheduling.
// ---------------------------------------------------------------------------
struct PopBlock {
  int32_t sdom;
  int32_t reserved;
  int64_t begin, end;  // element range in the subdomain chunk (memory order)
};

__global__ __launch_bounds__(256) void k_pop_partial(const DevSdom *__restrict__ sd, const PopBlock *__restrict__ blocks,
                                                     double *__restrict__ partial) {
  __shared__ double red[256];
  const PopBlock B = blocks[blockIdx.x];
  const DevSdom &S = sd[B.sdom];
  // memory order of psi: roles sorted by decreasing stride
  const int64_t st[3] = {S.fa, S.fg, S.fz};
  const int32_t ex[3] = {S.nd, S.ng, S.nz};
  int o0 = 0, o1 = 1, o2 = 2;  // slow, mid, fast
  if (st[o0] < st[o1]) { int t = o0; o0 = o1; o1 = t; }
  if (st[o1] < st[o2]) { int t = o1; o1 = o2; o2 = t; }
  if (st[o0] < st[o1]) { int t = o0; o0 = o1; o1 = t; }
  const uint32_t nmid = ex[o1], nfast = ex[o2];
  double acc = 0.0;
  for (int64_t e = B.begin + threadIdx.x; e < B.end; e += blockDim.x) {
    // exact: kr_problem_create rejects subdomain chunks above 2^31 elements
    const uint32_t ue = (uint32_t)e;
    const uint32_t f = ue % nfast, t = ue / nfast;
    const uint32_t m = t % nmid, s = t / nmid;
    uint32_t ix[3];
    ix[o0] = s; ix[o1] = m; ix[o2] = f;
    const double v = __builtin_nontemporal_load(S.psi + e);
    acc += S.w[ix[0]] * v * S.volume[ix[2]];
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(256) void k_pop_final(const double *__restrict__ partial, int64_t n,
                                                   double *__restrict__ out) {
  __shared__ double red[256];
  double acc = 0.0;
  for (int64_t e = threadIdx.x; e < n; e += blockDim.x) acc += partial[e];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0];
}

// ---------------------------------------------------------------------------
// kConst / kCopy: src/Kripke/Kernel.h:37-81.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_fill(double *__restrict__ p, int64_t n, double v) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
    p[e] = v;
}

__global__ __launch_bounds__(256) void k_copy(double *__restrict__ d, const double *__restrict__ s, int64_t n) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
    d[e] = s[e];
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------

#define KR_MOMENT_CASES(X) X(1) X(4) X(9) X(16) X(25) X(36) X(49) X(64)

// Directions per batch: the largest of 16, 8, 4, 2, 1 that divides nd.
static int batch_of(int nd, int cap) {
  int db = cap;
  while (db > 1 && nd % db != 0) db /= 2;
  return db;
}

// LDS tile row stride for the direction-fastest kernels: odd, so that a
// thread-per-row read of consecutive lanes hits distinct banks.  Tiles above
// the device's LDS per workgroup fall back to the column kernels.
static int dfast_stride(int nd, int M) { return std::max(nd, M) | 1; }
static bool dfast_fits(const LaunchCtx &L, int M) {
  return (size_t)DF_CB * dfast_stride(L.nd, M) * sizeof(double) <= L.lds_max;
}

template <int M, bool ACC, bool ADJ>
static void ltimes_launch(const LaunchCtx &L, dim3 grid, int db) {
  if (L.dfast && dfast_fits(L, M)) {
    const int ts = dfast_stride(L.nd, M);
    dim3 g((unsigned)((L.max_gz + DF_CB - 1) / DF_CB), L.num_pr);
    const size_t lds = (size_t)DF_CB * ts * sizeof(double);
    if (lds > 65536) allow_lds((const void *)k_ltimes_dfast<M, ACC, ADJ>, lds);
    hipLaunchKernelGGL((k_ltimes_dfast<M, ACC, ADJ>), g, dim3(DF_CB), lds, L.stream, L.sd, L.pr, L.pr_sdoms, ts);
    return;
  }
  switch (db) {
    case 16: hipLaunchKernelGGL((k_ltimes<M, 16, ACC, ADJ>), grid, dim3(128), 0, L.stream, L.sd, L.pr, L.pr_sdoms); break;
    case 8: hipLaunchKernelGGL((k_ltimes<M, 8, ACC, ADJ>), grid, dim3(128), 0, L.stream, L.sd, L.pr, L.pr_sdoms); break;
    case 4: hipLaunchKernelGGL((k_ltimes<M, 4, ACC, ADJ>), grid, dim3(128), 0, L.stream, L.sd, L.pr, L.pr_sdoms); break;
    case 2: hipLaunchKernelGGL((k_ltimes<M, 2, ACC, ADJ>), grid, dim3(128), 0, L.stream, L.sd, L.pr, L.pr_sdoms); break;
    default: hipLaunchKernelGGL((k_ltimes<M, 1, ACC, ADJ>), grid, dim3(128), 0, L.stream, L.sd, L.pr, L.pr_sdoms); break;
  }
}

template <int M, bool ACC, bool ADJ>
static void lplus_launch(const LaunchCtx &L, dim3 grid, int db) {
  if (L.dfast && dfast_fits(L, M)) {
    const int ts = dfast_stride(L.nd, M);
    dim3 g((unsigned)((L.max_gz + DF_CB - 1) / DF_CB), L.num_pr);
    const size_t lds = (size_t)DF_CB * ts * sizeof(double);
    if (lds > 65536) allow_lds((const void *)k_lplustimes_dfast<M, ACC, ADJ>, lds);
    hipLaunchKernelGGL((k_lplustimes_dfast<M, ACC, ADJ>), g, dim3(DF_CB), lds, L.stream, L.sd, L.pr, L.pr_sdoms, ts);
    return;
  }
  switch (db) {
    case 4: hipLaunchKernelGGL((k_lplustimes<M, 4, ACC, ADJ>), grid, dim3(256), 0, L.stream, L.sd, L.pr, L.pr_sdoms); break;
    case 2: hipLaunchKernelGGL((k_lplustimes<M, 2, ACC, ADJ>), grid, dim3(256), 0, L.stream, L.sd, L.pr, L.pr_sdoms); break;
    default: hipLaunchKernelGGL((k_lplustimes<M, 1, ACC, ADJ>), grid, dim3(256), 0, L.stream, L.sd, L.pr, L.pr_sdoms); break;
  }
}

// mode: 0 overwrite, 1 accumulate, 2 adjoint (overwrite)
static int ltimes_shape(const LaunchCtx &L, int mode, const char *what, bool lplus) {
  // LTimes' column kernel runs 128-thread blocks, LPlusTimes' 256: on 64^3
  // zones x 128 groups x 128 directions at P4 (DGZ), 128 took LTimes from
  // 7.3 to 6.95 ms (mean of three runs) and LPlusTimes 0.1-0.9 ms slower
  dim3 grid(grid_for(L.max_gz, lplus ? 256 : 128, 1 << 30), L.num_pr);
  // LPlusTimes writes 2 directions per batch: on 64^3 zones x 128 groups x
  // 128 directions at P4 the kernel took 10.5 ms with 8 in flight per thread,
  // about 7.9 ms with 4 and about 7.4 ms with 1 or 2
  const int db = batch_of(L.nd, lplus ? 2 : 16);
  switch (L.num_moments) {
#define X(m)                                                     \
  case m:                                                        \
    if (lplus) {                                                 \
      if (mode == 1) lplus_launch<m, true, false>(L, grid, db);  \
      else if (mode == 0) lplus_launch<m, false, false>(L, grid, db); \
      else lplus_launch<m, false, true>(L, grid, db);            \
    } else {                                                     \
      if (mode == 1) ltimes_launch<m, true, false>(L, grid, db); \
      else if (mode == 0) ltimes_launch<m, false, false>(L, grid, db); \
      else ltimes_launch<m, false, true>(L, grid, db);           \
    }                                                            \
    break;
    KR_MOMENT_CASES(X)
#undef X
    default: set_error(std::string(what) + ": unsupported number of moments"); return KR_ERR_UNSUPPORTED;
  }
  KR_HIP_CHECK(hipGetLastError());
  return KR_OK;
}

int64_t ltimes_variant(const LaunchCtx &L, bool lplus) {
  if (L.dfast && dfast_fits(L, L.num_moments)) return ((int64_t)L.num_moments << 32) | (1LL << 31) | dfast_stride(L.nd, L.num_moments);
  return ((int64_t)L.num_moments << 32) | batch_of(L.nd, lplus ? 2 : 16);
}

int launch_ltimes(const LaunchCtx &L, int accumulate) { return ltimes_shape(L, accumulate ? 1 : 0, "LTimes", false); }
int launch_lplustimes(const LaunchCtx &L, int accumulate) {
  return ltimes_shape(L, accumulate ? 1 : 0, "LPlusTimes", true);
}
// psi = L^T phi (the LPlusTimes contraction with ell)
int launch_ltimes_adjoint(const LaunchCtx &L) { return ltimes_shape(L, 2, "LTimes adjoint", true); }
// phi_out = (L+)^T rhs (the LTimes contraction with ell_plus)
int launch_lplustimes_adjoint(const LaunchCtx &L) { return ltimes_shape(L, 2, "LPlusTimes adjoint", false); }

int launch_scattering(const LaunchCtx &L, const int32_t *pr_gset, const int32_t *pr_r, int accumulate, bool adjoint) {
  SigsStrides ss{L.sigs_str[0], L.sigs_str[1], L.sigs_str[2], L.sigs_str[3]};
  // 1) classify data/sigs as it is now; 2) the diagonal path and the dense
  // path are both enqueued and each returns at once unless it applies.
  hipLaunchKernelGGL(k_sigs_structure, dim3(SIGS_FLAG_BLOCKS), dim3(256), 0, L.stream, L.sigs, ss, L.sigs_dims[0],
                     L.sigs_dims[1], L.sigs_dims[2], L.sigs_flags);
  KR_HIP_CHECK(hipGetLastError());
  dim3 grid(grid_for(L.max_gz), L.num_pr);
  const size_t tile_lds = (size_t)DF_CB * (L.num_moments | 1) * sizeof(double);
  if (L.dfast && !accumulate && tile_lds <= L.lds_max) {
    // phi_out is [column][nm] in these nestings: coalesced through an LDS tile
    const dim3 g2((unsigned)((L.max_gz + DF_CB - 1) / DF_CB), L.num_pr);
    int rc;
    if (tile_lds > 65536 && ((rc = allow_lds((const void *)k_scatter_diag_dfast<true>, tile_lds)) ||
                             (rc = allow_lds((const void *)k_scatter_diag_dfast<false>, tile_lds))))
      return rc;
    if (adjoint)
      hipLaunchKernelGGL(k_scatter_diag_dfast<true>, g2, dim3(DF_CB), tile_lds, L.stream, L.pr, L.r, L.sigs, ss,
                         L.num_moments, (const int32_t *)L.sigs_flags);
    else
      hipLaunchKernelGGL(k_scatter_diag_dfast<false>, g2, dim3(DF_CB), tile_lds, L.stream, L.pr, L.r, L.sigs, ss,
                         L.num_moments, (const int32_t *)L.sigs_flags);
  } else if (adjoint) {
    hipLaunchKernelGGL(k_scatter_diag<true>, grid, dim3(256), 0, L.stream, L.pr, L.r, L.sigs, ss, L.num_moments,
                       accumulate, (const int32_t *)L.sigs_flags);
  } else {
    hipLaunchKernelGGL(k_scatter_diag<false>, grid, dim3(256), 0, L.stream, L.pr, L.r, L.sigs, ss, L.num_moments,
                       accumulate, (const int32_t *)L.sigs_flags);
  }
  KR_HIP_CHECK(hipGetLastError());
  // Dense: coefficient re-layout, then for each source chunk (work-list
  // order) and each destination chunk on the same zone set
  // (Scattering.cpp:137-145).  The transpose writes the source chunks: the
  // first pair that writes a chunk overwrites (unless accumulating), later
  // pairs add.
  const int64_t nmat = L.sigs_dims[0], L1 = L.sigs_dims[1], G = L.sigs_dims[2];
  const int M = L.num_moments;
  const int32_t *flags = (const int32_t *)L.sigs_flags;
  hipLaunchKernelGGL(k_sigs_relayout, dim3(grid_for(nmat * L1 * G * G + SG_PAD)), dim3(256), 0, L.stream, L.sigs, ss,
                     nmat, L1, G, L.sigs_t, adjoint ? 1 : 0, flags);
  KR_HIP_CHECK(hipGetLastError());
  int32_t *tab = (int32_t *)(L.sigs_t + nmat * L1 * G * G + SG_PAD);
  const int64_t wstride = L.max_z + GF_ZT;
  double *W = L.sigs_t + nmat * L1 * G * G + SG_PAD + tab_doubles(L1, M);
  hipLaunchKernelGGL(k_mat_weights, dim3(grid_for(L.max_z), L.num_r), dim3(256), 0, L.stream, L.r, nmat, wstride, W,
                     flags);
  KR_HIP_CHECK(hipGetLastError());
  if (L.dfast) {
    hipLaunchKernelGGL(k_moment_orders, dim3(1), dim3(64), 0, L.stream, L.m2l, M, (int)L1, tab, flags);
    KR_HIP_CHECK(hipGetLastError());
  }
  std::vector<int> first(L.num_pr, 1);
  for (int s = 0; s < L.num_pr; ++s)
    for (int d = 0; d < L.num_pr; ++d) {
      if (pr_r[s] != pr_r[d]) continue;
      const int w = adjoint ? s : d;
      const int acc = first[w] ? accumulate : 1;
      first[w] = 0;
      if (L.dfast) {
        const int nocg = (int)((L.max_g + SG_OC - 1) / SG_OC);
        const int64_t nst = (L.max_z + MF_ZS - 1) / MF_ZS;
        const int64_t nblk = ((nst + 7) / 8) * 8 * L1 * nocg;
        const unsigned gx = (unsigned)std::min<int64_t>(nblk, DENSE_MAX_BLOCKS_MFAST);  // both multiples of 8
        if (adjoint)
          hipLaunchKernelGGL(k_scatter_dense_mfast<true>, dim3(gx), dim3(256), 0, L.stream, L.pr, W, wstride,
                             L.sigs_t, tab, nmat, L1, G, s, d, acc, flags, nocg, nblk);
        else
          hipLaunchKernelGGL(k_scatter_dense_mfast<false>, dim3(gx), dim3(256), 0, L.stream, L.pr, W, wstride,
                             L.sigs_t, tab, nmat, L1, G, s, d, acc, flags, nocg, nblk);
      } else if (L.gfast) {
        const int64_t tiles = (L.max_z + GF_ZT - 1) / GF_ZT;
        const int64_t gxf = std::max<int64_t>(1, std::min<int64_t>((tiles + 3) / 4, 1 << 20));
        const int64_t gyf = (int64_t)M * ((L.max_g + 63) / 64);
        dim3 grid((unsigned)gxf, (unsigned)std::max<int64_t>(1, std::min<int64_t>(gyf, DENSE_MAX_BLOCKS / gxf)));
        if (adjoint)
          hipLaunchKernelGGL(k_scatter_dense_gfast<true>, grid, dim3(256), 0, L.stream, L.pr, W, wstride, L.sigs_t, L.m2l,
                             M, nmat, L1, G, s, d, acc, flags);
        else
          hipLaunchKernelGGL(k_scatter_dense_gfast<false>, grid, dim3(256), 0, L.stream, L.pr, W, wstride, L.sigs_t, L.m2l,
                             M, nmat, L1, G, s, d, acc, flags);
      } else {
        const int64_t gxf = (L.max_z + 255) / 256;
        const int64_t gyf = (int64_t)M * ((L.max_g + SG_OC - 1) / SG_OC);
        dim3 grid((unsigned)gxf, (unsigned)std::max<int64_t>(1, std::min<int64_t>(gyf, DENSE_MAX_BLOCKS / gxf)));
        if (adjoint)
          hipLaunchKernelGGL(k_scatter_dense<true>, grid, dim3(256), 0, L.stream, L.pr, W, wstride, L.sigs_t, L.m2l, M,
                             nmat, L1, G, s, d, acc, flags);
        else
          hipLaunchKernelGGL(k_scatter_dense<false>, grid, dim3(256), 0, L.stream, L.pr, W, wstride, L.sigs_t, L.m2l, M,
                             nmat, L1, G, s, d, acc, flags);
      }
      KR_HIP_CHECK(hipGetLastError());
    }
  (void)pr_gset;
  return KR_OK;
}

// Dense-path scratch: T (+ SG_PAD), the moment-order table (int32), then the
// weight table W[r][mat][z] with row stride max_z + GF_ZT.
int64_t sigs_scratch_doubles(int64_t nmat, int64_t L1, int64_t G, int64_t M, int64_t nr, int64_t max_z) {
  return nmat * L1 * G * G + SG_PAD + tab_doubles(L1, M) + nr * nmat * (max_z + GF_ZT);
}

int launch_scattering_zonal(const LaunchCtx &L, const double *const *sz_dev, int64_t s_g, int64_t s_z, int accumulate,
                            bool adjoint) {
  dim3 grid(grid_for(L.max_gz), L.num_pr);
  if (adjoint)
    hipLaunchKernelGGL(k_scatter_zonal<true>, grid, dim3(256), 0, L.stream, L.pr, sz_dev, s_g, s_z, L.num_moments,
                       accumulate);
  else
    hipLaunchKernelGGL(k_scatter_zonal<false>, grid, dim3(256), 0, L.stream, L.pr, sz_dev, s_g, s_z, L.num_moments,
                       accumulate);
  KR_HIP_CHECK(hipGetLastError());
  return KR_OK;
}

int launch_sigs_structure_multi(hipStream_t st, const DevScat *probs, int nprob) {
  if (nprob <= 0) return KR_OK;
  hipLaunchKernelGGL(k_sigs_structure_multi, dim3(SIGS_FLAG_BLOCKS, nprob), dim3(256), 0, st, probs);
  KR_HIP_CHECK(hipGetLastError());
  return KR_OK;
}

int launch_scatter_diag_multi(hipStream_t st, const DevPR *pr, const DevR *r, const DevScat *chunks, int nchunks,
                              int64_t max_gz) {
  if (nchunks <= 0) return KR_OK;
  hipLaunchKernelGGL(k_scatter_diag_multi, dim3(grid_for(max_gz), nchunks), dim3(256), 0, st, pr, r, chunks);
  KR_HIP_CHECK(hipGetLastError());
  return KR_OK;
}

int launch_source_multi(hipStream_t st, const DevPR *pr, const DevR *r, const DevScat *chunks, int nchunks,
                        int64_t max_gmix) {
  if (nchunks <= 0) return KR_OK;
  hipLaunchKernelGGL(k_source_multi, dim3(grid_for(max_gmix), nchunks), dim3(256), 0, st, pr, r, chunks);
  KR_HIP_CHECK(hipGetLastError());
  return KR_OK;
}

int launch_pop_sum_multi(hipStream_t st, const double *partial, const int64_t *list, const int64_t *begin, int nprob,
                         double *out, int64_t out_stride) {
  if (nprob <= 0) return KR_OK;
  hipLaunchKernelGGL(k_pop_sum_multi, dim3(nprob), dim3(256), 0, st, partial, list, begin, out, out_stride);
  KR_HIP_CHECK(hipGetLastError());
  return KR_OK;
}

int launch_source(const LaunchCtx &L, double strength) {
  dim3 grid(grid_for(L.max_gmix), L.num_pr);
  hipLaunchKernelGGL(k_source, grid, dim3(256), 0, L.stream, L.pr, L.r, strength);
  KR_HIP_CHECK(hipGetLastError());
  return KR_OK;
}

int launch_population(const LaunchCtx &L, const void *blocks, int64_t num_blocks, double *partial, double *out) {
  hipLaunchKernelGGL(k_pop_partial, dim3((unsigned)num_blocks), dim3(256), 0, L.stream, L.sd,
                     (const PopBlock *)blocks, partial);
  KR_HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_pop_final, dim3(1), dim3(256), 0, L.stream, partial, num_blocks, out);
  KR_HIP_CHECK(hipGetLastError());
  return KR_OK;
}

int launch_vol_info(hipStream_t st, const DevR *r, int nr) {
  if (nr <= 0) return KR_OK;
  hipLaunchKernelGGL(k_vol_info, dim3(nr), dim3(1024), 0, st, r);
  KR_HIP_CHECK(hipGetLastError());
  return KR_OK;
}

int launch_sum(hipStream_t st, const double *partial, int64_t n, double *out) {
  hipLaunchKernelGGL(k_pop_final, dim3(1), dim3(256), 0, st, partial, n, out);
  KR_HIP_CHECK(hipGetLastError());
  return KR_OK;
}

int64_t population_block_table(const DevSdom *host_sd, int nsdom, std::vector<char> &bytes) {
  std::vector<PopBlock> b;
  const int64_t per_block = 256 * 16;
  for (int s = 0; s < nsdom; ++s) {
    const int64_t n = (int64_t)host_sd[s].nd * host_sd[s].ng * host_sd[s].nz;
    for (int64_t e = 0; e < n; e += per_block) b.push_back(PopBlock{s, 0, e, std::min(n, e + per_block)});
  }
  bytes.assign((const char *)b.data(), (const char *)(b.data() + b.size()));
  return (int64_t)b.size();
}

int launch_fill(hipStream_t st, double *p, int64_t n, double v) {
  if (n <= 0) return KR_OK;
  hipLaunchKernelGGL(k_fill, dim3(grid_for(n, 256, 65536)), dim3(256), 0, st, p, n, v);
  KR_HIP_CHECK(hipGetLastError());
  return KR_OK;
}

int launch_copy(hipStream_t st, double *d, const double *s, int64_t n) {
  if (n <= 0) return KR_OK;
  hipLaunchKernelGGL(k_copy, dim3(grid_for(n, 256, 65536)), dim3(256), 0, st, d, s, n);
  KR_HIP_CHECK(hipGetLastError());
  return KR_OK;
}

}  // namespace kr
