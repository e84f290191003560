// Registered state behind the C-ABI handles (kr_context, kr_problem),
// shared by the sources that implement the entry points.
#pragma once
#include <map>
#include <string>
#include <vector>

#include "kr_launch.h"

using kr::DevPR;
using kr::DevR;
using kr::DevSdom;
using kr::SweepPlanes;
using kr::SweepShape;
using kr::SweepUnit;

// Autotuner result for one problem shape (kr_problem_autotune).
struct TuneChoice {
  int v_max = 0, cap_div = 1;
  int r_scale_pct = 100;  // pair groups per work unit, percent of the default
  double ms = 0.0;
};

struct kr_context {
  int device;
  hipStream_t stream;
  size_t lds_max = 65536;  // LDS per workgroup (hipDeviceAttributeMaxSharedMemoryPerBlock)
  double *noalloc_scratch = nullptr;  // this device's address of the no-allocation scratch array
  std::map<std::vector<int64_t>, TuneChoice> tuned;  // shape key -> choice
};

struct TimedCall {
  std::string name;
  hipEvent_t start, stop;
};

struct kr_problem {
  kr_context *ctx = nullptr;
  int nesting = 0, nsdom = 0, npr = 0, nr = 0, M = 0, L1 = 0, nmat = 0, G = 0;
  std::vector<kr_sdom_desc> sdoms;
  std::vector<kr_pr_desc> prs;
  std::vector<kr_r_desc> rs;
  std::vector<DevSdom> h_sd;
  std::vector<DevPR> h_pr;
  std::vector<int32_t> pr_gset, pr_r;
  DevSdom *d_sd = nullptr;
  DevPR *d_pr = nullptr;
  DevR *d_r = nullptr;
  int32_t *d_pr_sdoms = nullptr;
  double *d_coef = nullptr;
  SweepPlanes *d_planes_ref = nullptr, *d_planes_pipe = nullptr;
  // sweep schedules
  SweepShape shape_ref{}, shape_pipe{};
  bool aligned16 = true;
  TuneChoice tune;
  int64_t sched_version = 0;  // bumped when the pipelined schedule is replaced (captured graphs go stale)
  SweepUnit *d_units_ref = nullptr;            // all sdoms, grouped by sdom
  std::vector<int32_t> ref_unit_begin, ref_unit_count;
  SweepUnit *d_units_pipe = nullptr;           // all sdoms, grouped by level
  std::vector<int32_t> pipe_level_begin, pipe_level_count;
  SweepUnit *d_units_tmp = nullptr;            // scratch for arbitrary lists
  int32_t units_tmp_cap = 0;
  double *d_pop_units = nullptr;               // fused population: one partial per sweep unit
  int64_t pop_units_cap = 0;
  SweepPlanes *d_planes_adj = nullptr;         // adjoint sweep: faces from the forward-downwind side
  SweepPlanes *d_planes_adj_seed = nullptr;    // the same, with inflow on the domain boundary from the own planes
  SweepPlanes *d_planes_bj = nullptr;          // Block Jacobi: no incoming faces, outgoing faces kept in place
  static constexpr int PTR_SLOTS = 4;          // zonal sigma_s, sigma_t gradient, sigma_s gradient, spare
  const double **d_ptrs = nullptr;             // PTR_SLOTS per-PR-chunk pointer arrays passed to kernels
  std::vector<const double *> ptrs_last[PTR_SLOTS];  // their current contents
  // population
  void *d_pop_blocks = nullptr;
  int64_t pop_nblocks = 0;
  double *d_pop_partial = nullptr, *d_pop_out = nullptr;
  // scattering
  const double *sigs = nullptr;
  const int64_t *m2l = nullptr;
  int64_t sigs_str[4] = {0, 0, 0, 0};
  int32_t *d_sigs_flags = nullptr;
  double *d_sigs_t = nullptr;                 // dense scattering: sigs re-laid output-group fastest
  double *d_vol_info = nullptr;               // per R chunk: {uniform volume flag, volume} (k_vol_info)
  int64_t max_gz = 0, max_gmix = 0, max_g = 0, max_z = 0;
  // timing
  int timing = 0;
  std::vector<TimedCall> pending;
  std::map<std::string, std::pair<double, int64_t>> timers;
  // batches this problem belongs to (kr_batch_create); destroying the
  // problem marks them unusable
  std::vector<kr_batch *> batches;
};
