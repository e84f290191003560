// Internal device-side tables and helpers shared by the HIP sources.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "../../include/kripke_amd.h"

namespace kr {

// Pointers into HBM as the global address space: loads and stores become
// global_* instructions (counted on vmcnt only), not flat ones, so LDS waits
// do not drain in-flight HBM traffic.
template <class T>
using gmem = __attribute__((address_space(1))) T;

template <class T>
__device__ __forceinline__ const gmem<T> *gp(const T *p) {
  return (const gmem<T> *)p;
}
template <class T>
__device__ __forceinline__ gmem<T> *gp(T *p) {
  return (gmem<T> *)p;
}

// Read-only tables that stay constant for a whole launch (quadrature,
// ell/ell_plus): wave-uniform reads through this address space become scalar
// loads into SGPRs, which VALU instructions take as operands directly.
template <class T>
using cmem = __attribute__((address_space(4))) T;

template <class T>
__device__ __forceinline__ const cmem<T> *cp(const T *p) {
  return (const cmem<T> *)p;
}

// Strides of a (A, Group, Zone) field, A = Direction or Moment, for a
// nesting (src/Kripke/VarTypes.h:68-101).  Plane fields use the same rule
// with the face index (slow*nfast + fast) in place of the zone.
struct Str3 {
  int64_t a, g, z;
};

__host__ __device__ inline Str3 flux_strides(int nesting, int64_t na, int64_t ng, int64_t nz) {
  switch (nesting) {
    case KR_DGZ: return {ng * nz, nz, 1};
    case KR_DZG: return {nz * ng, 1, ng};
    case KR_GDZ: return {nz, na * nz, 1};
    case KR_GZD: return {1, nz * na, na};
    case KR_ZDG: return {ng, 1, na * ng};
    default: return {1, na, ng * na};  // KR_ZGD
  }
}

// Field_SigmaTZonal <Group, Zone>: [g][z] when Group precedes Zone.
__host__ __device__ inline void sigt_strides(int nesting, int64_t ng, int64_t nz, int64_t *sg, int64_t *sz) {
  if (nesting == KR_DGZ || nesting == KR_GDZ || nesting == KR_GZD) {
    *sg = nz;
    *sz = 1;
  } else {
    *sg = 1;
    *sz = ng;
  }
}

// Device copy of one subdomain's pointers, extents and strides.
struct DevSdom {
  double *psi, *rhs;
  double *plane[3];           // own i/j/k plane chunks
  const double *sig;          // sigt_zonal chunk of the subdomain's PR chunk
  double *coef[3];            // 2*cos(d)/delta tables: [d][i], [d][j], [d][k]
  const double *cosines[3];   // quadrature/xcos, ycos, zcos chunks
  const double *delta[3];     // dx, dy, dz chunks
  const double *w, *ell, *ell_plus;
  const double *volume;
  const double *vol_info;     // {1 if every zone volume of the R chunk is the same, that volume} (k_vol_info)
  int32_t nd, ng, ni, nj, nk, nz;
  int32_t id, jd, kd;
  int32_t pr, r, gset, dset, glo;
  int32_t up[3], down[3];
  int64_t fa, fg, fz;         // flux strides
  int64_t sg, sz;             // sigt strides
  int64_t pa[3], pg[3], pf[3];// plane strides (face index f)
};

// Device copy of one PR chunk: the moment fields and the subdomains that
// contribute to it (same group set and zone set), in SdomId order.
struct DevPR {
  double *phi, *phi_out;
  const double *sig;
  int32_t ng, nz, r, gset, glo;
  int32_t sdom_begin, sdom_count;  // range in the pr_sdoms list
  int64_t ma, mg, mz;              // moment-field strides
};

struct DevR {
  const int64_t *zone_to_mixelem;
  const int32_t *zone_to_num;
  const int64_t *mix_zone, *mix_mat;
  const double *mix_frac;
  const double *volume;
  double *vol_info;           // written by k_vol_info: {uniform flag, first volume}
  int64_t num_mix;
  int32_t nz;
};

// One sweep work unit: a run of pair groups of one subdomain.
struct SweepUnit {
  int32_t sdom;
  int32_t pg_begin;   // first pair group
  int32_t pg_count;   // pair groups in this unit
  int32_t reserved;
};

// Face sources and destinations of one subdomain for one sweep mode.
struct SweepPlanes {
  const double *in[3];  // incoming face per axis (nullptr = vacuum, reads as zero)
  double *out[3];       // destination of faces leaving the subdomain
};

// Per-launch sweep geometry (same for every unit of a launch).
struct SweepShape {
  int32_t kind;                 // 0: systolic pencil kernel, 1: row-scan kernel (zone-fastest nestings), 2: none fits,
                                // 3: wave-uniform plane kernel (Q = 64 pairs, PK = waves, NPK = k-patches)
  int32_t Q, PJ, PK, NPJ, NPK;  // lanes per zone column, patch extents, patches per axis
  int32_t RG;                   // row-scan kernel: lanes per row group
  int32_t V;                    // row-scan kernel: consecutive zones per lane
  int32_t inner_is_g;           // pair index = outer*inner + q ordering
  int32_t ref_mode;             // 1: read/write own planes only (Kernel::sweepSubdomain)
  int32_t nd, ni, nj, nk;       // subdomain extents
  int32_t kring;                // systolic kernel, patches in k-fastest order: faces between k-patches in an LDS ring
};

void set_error(const std::string &msg);

}  // namespace kr

#define KR_HIP_CHECK(expr)                                                                    \
  do {                                                                                        \
    hipError_t _e = (expr);                                                                   \
    if (_e != hipSuccess) {                                                                   \
      kr::set_error(std::string(#expr) + ": " + hipGetErrorString(_e));                       \
      return KR_ERR_HIP;                                                                      \
    }                                                                                         \
  } while (0)
