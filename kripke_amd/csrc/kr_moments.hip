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
// A thread owns one (g,z) column: it issues the psi loads of a batch of DB
// directions, then folds them into M accumulators, direction by direction,
// with ell read through the scalar cache (SGPR operands of the FMAs).
// ---------------------------------------------------------------------------
// DB directions of one column: all psi loads first, then the FMAs.
template <int M, int DB>
__device__ __forceinline__ void ltimes_batch(double (&acc)[M], const gmem<double> *psi, int64_t fa,
                                             const cmem<double> *ell) {
  double p[DB];
#pragma unroll
  for (int b = 0; b < DB; ++b) p[b] = __builtin_nontemporal_load(psi + b * fa);
  __builtin_amdgcn_sched_barrier(0);  // keep all DB loads in flight before the first FMA
#pragma unroll
  for (int b = 0; b < DB; ++b) {
#pragma unroll
    for (int nm = 0; nm < M; ++nm) acc[nm] = __fma_rn(ell[b * M + nm], p[b], acc[nm]);
  }
}

template <int M, int DB, bool ACC>
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
  gmem<double> *phi = gp(P.phi);
  double acc[M];
#pragma unroll
  for (int nm = 0; nm < M; ++nm) acc[nm] = ACC ? phi[po + nm * P.ma] : 0.0;
  for (int k = 0; k < P.sdom_count; ++k) {
    const DevSdom &S = sd[pr_sdoms[P.sdom_begin + k]];
    const gmem<double> *psi = gp(S.psi) + g * S0.fg + z * S0.fz;
    const cmem<double> *ell = cp(S.ell);
    for (int d0 = 0; d0 < nd; d0 += DB) ltimes_batch<M, DB>(acc, psi + d0 * fa, fa, ell + d0 * M);  // DB divides nd
  }
#pragma unroll
  for (int nm = 0; nm < M; ++nm) phi[po + nm * P.ma] = acc[nm];
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

template <int M, int DB, bool ACC>
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
  const gmem<double> *pout = gp(P.phi_out);
  double pho[M];
#pragma unroll
  for (int nm = 0; nm < M; ++nm) pho[nm] = pout[po + nm * P.ma];
  for (int k = 0; k < P.sdom_count; ++k) {
    const DevSdom &S = sd[pr_sdoms[P.sdom_begin + k]];
    gmem<double> *rhs = gp(S.rhs) + g * S0.fg + z * S0.fz;
    const cmem<double> *elp = cp(S.ell_plus);
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
constexpr int DF_CB = 128;  // columns per block

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

__device__ __forceinline__ void tile_to_slab(gmem<double> *dst, const double *tile, int ts, int total, int w) {
  SlabWalk sw(threadIdx.x, w);
  for (int e = threadIdx.x; e < total; e += DF_CB) {
    __builtin_nontemporal_store(tile[sw.col * ts + sw.k], dst + e);
    sw.next(w);
  }
}

template <int M, bool ACC>
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
  gmem<double> *phi = gp(P.phi) + c0 * M;
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
  for (int k = 0; k < P.sdom_count; ++k) {
    const DevSdom &S = sd[pr_sdoms[P.sdom_begin + k]];
    slab_to_tile<true>(tile, ts, gp(S.psi) + c0 * nd, cn * nd, nd);
    __syncthreads();
    const cmem<double> *ell = cp(S.ell);
    if (tid < cn) {
      for (int d = 0; d < nd; ++d) {
        const double p = tile[tid * ts + d];
#pragma unroll
        for (int nm = 0; nm < M; ++nm) acc[nm] = __fma_rn(ell[d * M + nm], p, acc[nm]);
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int nm = 0; nm < M; ++nm) tile[tid * ts + nm] = acc[nm];
  __syncthreads();
  tile_to_slab(phi, tile, ts, cn * M, M);
}

template <int M, bool ACC>
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
  slab_to_tile<false>(tile, ts, gp(P.phi_out) + c0 * M, cn * M, M);
  __syncthreads();
  double pho[M];
#pragma unroll
  for (int nm = 0; nm < M; ++nm) pho[nm] = tid < cn ? tile[tid * ts + nm] : 0.0;
  __syncthreads();
  for (int k = 0; k < P.sdom_count; ++k) {
    const DevSdom &S = sd[pr_sdoms[P.sdom_begin + k]];
    gmem<double> *rhs = gp(S.rhs) + c0 * nd;
    if (ACC) {
      slab_to_tile<false>(tile, ts, rhs, cn * nd, nd);
      __syncthreads();
    }
    const cmem<double> *elp = cp(S.ell_plus);
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
    const double base0 = accumulate ? P.phi_out[po] : 0.0;
    const double t = s * P.phi[po];
    P.phi_out[po] = base0 + t;
    for (int nm = 1; nm < M; ++nm) {
      const double b = accumulate ? P.phi_out[po + nm * P.ma] : 0.0;
      P.phi_out[po + nm * P.ma] = b + 0.0;
    }
  }
}

// General sigs (group coupling, anisotropic moments): one thread per
// (moment, group, zone) of the destination chunk, source groups in order.
__global__ __launch_bounds__(256) void k_scatter_dense(const DevPR *__restrict__ prs, const DevR *__restrict__ rs,
                                                       const double *__restrict__ sigs, SigsStrides ss,
                                                       const int64_t *__restrict__ m2l, int M, int src, int dst,
                                                       int accumulate, const int32_t *__restrict__ flags) {
  if (!sigs_coupled(flags)) return;
  const DevPR &Ps = prs[src];
  const DevPR &Pd = prs[dst];
  const DevR &R = rs[Pd.r];
  const int64_t n = (int64_t)M * Pd.ng * Pd.nz;
  const bool zfast = Pd.mz <= Pd.mg;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t per = (int64_t)Pd.ng * Pd.nz;
    const int nm = (int)(e / per);
    int32_t g, z;
    gz_of(e % per, Pd.ng, Pd.nz, zfast, g, z);
    const int64_t nl = m2l[nm];
    const int64_t od = nm * Pd.ma + g * Pd.mg + z * Pd.mz;
    double acc = accumulate ? Pd.phi_out[od] : 0.0;
    for (int32_t gp = 0; gp < Ps.ng; ++gp) {
      const double s = sigs_zone(sigs, ss, R, z, nl, Pd.glo + g, Ps.glo + gp);
      const double t = s * Ps.phi[nm * Ps.ma + gp * Ps.mg + z * Ps.mz];
      acc = acc + t;
    }
    Pd.phi_out[od] = acc;
  }
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
static inline int grid_for(int64_t n, int per_block = 256, int max_blocks = 8192) {
  int64_t b = (n + per_block - 1) / per_block;
  if (b > max_blocks) b = max_blocks;
  if (b < 1) b = 1;
  return (int)b;
}

#define KR_MOMENT_CASES(X) X(1) X(4) X(9) X(16) X(25) X(36) X(49) X(64)

// Directions per batch: the largest of 16, 8, 4, 2, 1 that divides nd.
static int batch_of(int nd, int cap) {
  int db = cap;
  while (db > 1 && nd % db != 0) db /= 2;
  return db;
}

// LDS tile row stride for the direction-fastest kernels: odd, so that a
// thread-per-row read of consecutive lanes hits distinct banks.  Tiles above
// 64 KiB (very many directions or moments) fall back to the column kernels.
static int dfast_stride(int nd, int M) { return std::max(nd, M) | 1; }
static bool dfast_fits(int nd, int M) { return (size_t)DF_CB * dfast_stride(nd, M) * sizeof(double) <= 65536; }

template <int M, bool ACC>
static void ltimes_launch(const LaunchCtx &L, dim3 grid, int db) {
  if (L.dfast && dfast_fits(L.nd, M)) {
    const int ts = dfast_stride(L.nd, M);
    dim3 g((unsigned)((L.max_gz + DF_CB - 1) / DF_CB), L.num_pr);
    hipLaunchKernelGGL((k_ltimes_dfast<M, ACC>), g, dim3(DF_CB), (size_t)DF_CB * ts * sizeof(double), L.stream,
                       L.sd, L.pr, L.pr_sdoms, ts);
    return;
  }
  switch (db) {
    case 16: hipLaunchKernelGGL((k_ltimes<M, 16, ACC>), grid, dim3(256), 0, L.stream, L.sd, L.pr, L.pr_sdoms); break;
    case 8: hipLaunchKernelGGL((k_ltimes<M, 8, ACC>), grid, dim3(256), 0, L.stream, L.sd, L.pr, L.pr_sdoms); break;
    case 4: hipLaunchKernelGGL((k_ltimes<M, 4, ACC>), grid, dim3(256), 0, L.stream, L.sd, L.pr, L.pr_sdoms); break;
    case 2: hipLaunchKernelGGL((k_ltimes<M, 2, ACC>), grid, dim3(256), 0, L.stream, L.sd, L.pr, L.pr_sdoms); break;
    default: hipLaunchKernelGGL((k_ltimes<M, 1, ACC>), grid, dim3(256), 0, L.stream, L.sd, L.pr, L.pr_sdoms); break;
  }
}

template <int M, bool ACC>
static void lplus_launch(const LaunchCtx &L, dim3 grid, int db) {
  if (L.dfast && dfast_fits(L.nd, M)) {
    const int ts = dfast_stride(L.nd, M);
    dim3 g((unsigned)((L.max_gz + DF_CB - 1) / DF_CB), L.num_pr);
    hipLaunchKernelGGL((k_lplustimes_dfast<M, ACC>), g, dim3(DF_CB), (size_t)DF_CB * ts * sizeof(double), L.stream,
                       L.sd, L.pr, L.pr_sdoms, ts);
    return;
  }
  switch (db) {
    case 4: hipLaunchKernelGGL((k_lplustimes<M, 4, ACC>), grid, dim3(256), 0, L.stream, L.sd, L.pr, L.pr_sdoms); break;
    case 2: hipLaunchKernelGGL((k_lplustimes<M, 2, ACC>), grid, dim3(256), 0, L.stream, L.sd, L.pr, L.pr_sdoms); break;
    default: hipLaunchKernelGGL((k_lplustimes<M, 1, ACC>), grid, dim3(256), 0, L.stream, L.sd, L.pr, L.pr_sdoms); break;
  }
}

int launch_ltimes(const LaunchCtx &L, int accumulate) {
  dim3 grid(grid_for(L.max_gz, 256, 1 << 30), L.num_pr);
  const int db = batch_of(L.nd, 16);
  switch (L.num_moments) {
#define X(m)                                          \
  case m:                                             \
    if (accumulate) ltimes_launch<m, true>(L, grid, db); \
    else ltimes_launch<m, false>(L, grid, db);          \
    break;
    KR_MOMENT_CASES(X)
#undef X
    default: set_error("LTimes: unsupported number of moments"); return KR_ERR_UNSUPPORTED;
  }
  KR_HIP_CHECK(hipGetLastError());
  return KR_OK;
}

int launch_lplustimes(const LaunchCtx &L, int accumulate) {
  dim3 grid(grid_for(L.max_gz, 256, 1 << 30), L.num_pr);
  const int db = batch_of(L.nd, 4);
  switch (L.num_moments) {
#define X(m)                                         \
  case m:                                            \
    if (accumulate) lplus_launch<m, true>(L, grid, db); \
    else lplus_launch<m, false>(L, grid, db);          \
    break;
    KR_MOMENT_CASES(X)
#undef X
    default: set_error("LPlusTimes: unsupported number of moments"); return KR_ERR_UNSUPPORTED;
  }
  KR_HIP_CHECK(hipGetLastError());
  return KR_OK;
}

int launch_scattering(const LaunchCtx &L, const int32_t *pr_gset, const int32_t *pr_r, int accumulate) {
  SigsStrides ss{L.sigs_str[0], L.sigs_str[1], L.sigs_str[2], L.sigs_str[3]};
  // 1) classify data/sigs as it is now; 2) the diagonal path and the dense
  // path are both enqueued and each returns at once unless it applies.
  hipLaunchKernelGGL(k_sigs_structure, dim3(SIGS_FLAG_BLOCKS), dim3(256), 0, L.stream, L.sigs, ss, L.sigs_dims[0],
                     L.sigs_dims[1], L.sigs_dims[2], L.sigs_flags);
  KR_HIP_CHECK(hipGetLastError());
  dim3 grid(grid_for(L.max_gz), L.num_pr);
  hipLaunchKernelGGL(k_scatter_diag, grid, dim3(256), 0, L.stream, L.pr, L.r, L.sigs, ss, L.num_moments, accumulate,
                     (const int32_t *)L.sigs_flags);
  KR_HIP_CHECK(hipGetLastError());
  // Dense: for each source chunk (work-list order) and each destination
  // chunk on the same zone set (Scattering.cpp:137-145).
  std::vector<int> first(L.num_pr, 1);
  for (int s = 0; s < L.num_pr; ++s)
    for (int d = 0; d < L.num_pr; ++d) {
      if (pr_r[s] != pr_r[d]) continue;
      const int acc = first[d] ? accumulate : 1;
      first[d] = 0;
      hipLaunchKernelGGL(k_scatter_dense, dim3(grid_for(L.max_gz * L.num_moments)), dim3(256), 0, L.stream, L.pr,
                         L.r, L.sigs, ss, L.m2l, L.num_moments, s, d, acc, (const int32_t *)L.sigs_flags);
      KR_HIP_CHECK(hipGetLastError());
    }
  (void)pr_gset;
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
