// Launcher interface between the C-ABI layer and the kernel sources.
#pragma once
#include <vector>

#include "kr_internal.h"

namespace kr {

// Everything a moment-space launch needs, resolved from a registered problem.
struct LaunchCtx {
  hipStream_t stream;
  const DevSdom *sd;        // device
  const DevPR *pr;          // device
  const DevR *r;            // device
  const int32_t *pr_sdoms;  // device
  const double *sigs;       // device
  const int64_t *m2l;       // device
  int64_t sigs_str[4];      // mat, n, g, gp strides of data/sigs
  int64_t sigs_dims[4];     // materials, legendre, groups, groups
  int32_t *sigs_flags;      // device, SIGS_FLAG_BLOCKS ints: nonzero = a coupling entry seen
  double *sigs_t;           // device scratch of sigs_scratch_doubles(): sigs re-laid for the dense path, moment orders
  int num_pr, num_moments, num_r;
  int nd;                   // directions per subdomain (the same for every subdomain)
  int dfast;                // direction-fastest nesting (GZD, ZGD): psi [column][d], phi [column][nm]
  int gfast;                // group-fastest moment fields (DZG, ZDG): phi [..][z][g] / [z][nm][g]
  size_t lds_max;           // LDS per workgroup on this device
  int64_t max_gz;           // max groups*zones over PR chunks
  int64_t max_gmix;         // max groups*mixed elements over PR chunks
  int64_t max_g, max_z;     // max groups and zones of one PR chunk
};

constexpr int SIGS_FLAG_BLOCKS = 64;

// Scattering data of one problem, one entry per problem or per PR chunk in
// the batched (multi-problem) launches.
struct DevScat {
  const double *sigs;    // data/sigs
  int64_t str[4];        // mat, n, g, gp strides
  int64_t nmat, L1, G;
  int32_t *flags;        // the problem's SIGS_FLAG_BLOCKS structure flags
  double strength;       // source strength
  int32_t M;             // moments
  int32_t problem;
};
int launch_sigs_structure_multi(hipStream_t st, const DevScat *probs, int nprob);
// the diagonal scattering (overwrite) of every chunk whose problem's sigs is uncoupled
int launch_scatter_diag_multi(hipStream_t st, const DevPR *pr, const DevR *r, const DevScat *chunks, int nchunks,
                              int64_t max_gz);
int launch_source_multi(hipStream_t st, const DevPR *pr, const DevR *r, const DevScat *chunks, int nchunks,
                        int64_t max_gmix);
// out[p*out_stride] = sum of partial[list[begin[p] .. begin[p+1])] (k_pop_final's order)
int launch_pop_sum_multi(hipStream_t st, const double *partial, const int64_t *list, const int64_t *begin, int nprob,
                         double *out, int64_t out_stride);

// Kernel variant LTimes (lplus: LPlusTimes) takes for L: launches whose
// contexts give the same value run the same kernel, and every PR chunk's
// result depends only on its own tables, so chunks of several problems with
// equal variants can share one launch (tables combined, nd of any of them).
int64_t ltimes_variant(const LaunchCtx &L, bool lplus);
int launch_ltimes(const LaunchCtx &L, int accumulate);
int launch_lplustimes(const LaunchCtx &L, int accumulate);
// doubles of LaunchCtx::sigs_t for the dense scattering path
int64_t sigs_scratch_doubles(int64_t nmat, int64_t L1, int64_t G, int64_t M, int64_t nr, int64_t max_z);
int launch_scattering(const LaunchCtx &L, const int32_t *pr_gset, const int32_t *pr_r, int accumulate,
                      bool adjoint = false);
// per-zone isotropic group-diagonal scattering from s(g,z) arrays (device array of one pointer per PR chunk)
int launch_scattering_zonal(const LaunchCtx &L, const double *const *sz_dev, int64_t s_g, int64_t s_z, int accumulate,
                            bool adjoint);
// adjoints of the moment maps: psi = L^T phi; phi_out = (L+)^T rhs
int launch_ltimes_adjoint(const LaunchCtx &L);
int launch_lplustimes_adjoint(const LaunchCtx &L);
// parameter gradients (F: forward problem; adj_*: the adjoint problem's device tables)
int launch_grad_sigt(const LaunchCtx &F, const DevSdom *adj_sd, double *const *grad_dev, int64_t s_g, int64_t s_z,
                     int accumulate);
int launch_grad_sigs_zonal(const LaunchCtx &F, const DevPR *adj_pr, double *const *grad_dev, int64_t s_g, int64_t s_z,
                           int accumulate);
int launch_source(const LaunchCtx &L, double strength);
int launch_population(const LaunchCtx &L, const void *blocks, int64_t num_blocks, double *partial, double *out);
int64_t population_block_table(const DevSdom *host_sd, int nsdom, std::vector<char> &bytes);
int launch_fill(hipStream_t st, double *p, int64_t n, double v);
int launch_copy(hipStream_t st, double *d, const double *s, int64_t n);

// Sweep: one launch sweeps every unit; units of one launch share `shape`.
// pop (row-scan kernel only): one population partial per unit, or nullptr.
// adjoint: the transposed sweep (reads psi, writes rhs, axes reversed; planes
// and units of the reversed dependency order).
int launch_sweep(hipStream_t st, const DevSdom *sd, const SweepPlanes *planes, const SweepUnit *units,
                 int32_t num_units, const SweepShape &shape, double *pop = nullptr, bool adjoint = false);
// Row-scan sweep of the zone-fastest nestings (kr_sweep_rows.hip): blocks of
// shape.RG x shape.PK threads.  Units may belong to subdomains of different
// extents (each a multiple of shape.V along i and of shape.PK along k, with
// at most shape.RG*shape.V zones along i); coef_doubles: the largest
// nd*(ni+nj+nk+1) over their subdomains (the LDS coefficient table).
int launch_sweep_rows(hipStream_t st, const DevSdom *sd, const SweepPlanes *planes, const SweepUnit *units,
                      int32_t num_units, const SweepShape &shape, int64_t coef_doubles, double *pop, bool adjoint);
// Lets `kernel` take `bytes` of dynamic LDS (needed above 64 KiB).
int allow_lds(const void *kernel, size_t bytes);
// vol_info of R chunks [0, nr): read by the row-scan sweep's fused population.
int launch_vol_info(hipStream_t st, const DevR *r, int nr);
// Sum of n doubles in a fixed order (one workgroup): out[0] = sum(partial[0..n)).
int launch_sum(hipStream_t st, const double *partial, int64_t n, double *out);
// Refresh the 2*cos/delta tables of subdomains [first, first+count).
int launch_sweep_coef(hipStream_t st, const DevSdom *sd, int first, int count);
// Shape choice for a subdomain extent and nesting (pair lanes, patch).
// Tuning knobs of the shape choice (the autotuner's search space):
// v_max caps the row-scan kernel's zones per lane (0: no cap); cap_div
// shrinks the systolic kernel's largest patch by that factor.
struct SweepTune {
  int v_max = 0;           // row-scan: zones per lane cap; other nestings: 2 = the wave kernel, 3 = the systolic
                           // kernel with k-fastest patches and an LDS k-face ring, where they fit
  int cap_div = 1;         // systolic (both forms): patch size divisor; wave kernel: waves per workgroup
                           // = 16 / cap_div
  size_t lds_max = 65536;  // the device's LDS per workgroup (shapes within 64 KiB are preferred)
};
// Wave-uniform plane sweep of the nestings whose fastest index is a group or
// a direction (kr_sweep_waves.hip): 64 pairs per wave, one k-layer per wave,
// at most kt_max waves per workgroup.  Returns false when the subdomain does
// not fit it (pairs not a multiple of 64, ni not 8, 16 or 32).
bool waves_shape(SweepShape &sh, int32_t nd, int32_t ng, int32_t ni, int32_t nj, int32_t nk, size_t lds_limit,
                 int kt_max);
int launch_sweep_waves(hipStream_t st, const DevSdom *sd, const SweepPlanes *planes, const SweepUnit *units,
                       int32_t num_units, const SweepShape &shape, double *pop, bool adjoint);
// LDS bytes one sweep workgroup of `shape` takes, with coef_doubles of
// coefficient records (row-scan and wave kernels; the systolic kernel keeps
// a fixed record per direction).
size_t sweep_lds_bytes(const SweepShape &shape, int64_t coef_doubles);
// aligned16: every psi/rhs/sigma_t/plane chunk starts on a 16-byte boundary.
SweepShape choose_sweep_shape(int nesting, int32_t nd, int32_t ng, int32_t ni, int32_t nj, int32_t nk, int ref_mode,
                              bool aligned16, const SweepTune &tune = SweepTune());

}  // namespace kr
