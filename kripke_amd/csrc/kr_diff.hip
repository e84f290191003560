// Differentiable source iteration behind one C-ABI forward/backward pair
// (kr_diff_* in include/kripke_amd.h).  No counterpart in the reference: it
// differentiates K iterations of SteadyStateSolver.cpp:44-104 from psi = 0,
//
//     phi_n     = L psi_n
//     phi_out_n = S(s) phi_n + q          per-zone isotropic scattering s(g,z)
//     rhs_n     = L+ phi_out_n
//     psi_{n+1} = H(sigt)^-1 rhs_n        the diamond-difference sweep
//     phi_K     = L psi_K                 the output,
//
// with respect to sigt(g,z) (the sigt_zonal field) and s(g,z), one array per
// PR chunk in sigt_zonal's layout.  psi_{n+1} depends on psi_n only through
// phi_n, so a checkpoint is the small phi_n.  Backward, for a cotangent
// phibar_K of phi_K (and optionally of the last sweep's exiting faces):
//
//     psibar_K = L^T phibar_K
//     for n = K-1 .. 0:   phi_n and psi_{n+1} recomputed (each checkpoint
//                         interval's phi once from its checkpoint, then
//                         psi_{n+1} from phi_n)
//         rhsbar_n     = H^-T psibar_{n+1}       adjoint sweep (n = K-1: the
//                        faces' cotangent as adjoint inflow)
//         d sigt      += -sum_d rhsbar_n * psi_{n+1}
//         phi_outbar_n = (L+)^T rhsbar_n
//         d s         += phi_outbar_n(0) * phi_n(0)
//         phibar_n     = S^T phi_outbar_n
//         psibar_n     = L^T phibar_n
//
// Every step is one of the library's entry points on the context stream; the
// adjoint variables live in the adjoint problem's fields (each in the field
// of its primal variable), so no field is copied in the backward chain.
// Several parameter sets ("fields") share the two problems: each keeps its
// own copy of sigt and s and its own phi checkpoints, and one call runs any
// contiguous range of them.
#include <algorithm>
#include <string>
#include <vector>

#include "kr_state.h"

using namespace kr;

struct kr_diff {
  kr_problem *fwd = nullptr, *adj = nullptr;
  int K = 0, ck = 0, nck = 0, F = 0, npr = 0, nsdom = 0;
  double strength = 1.0;
  std::vector<int64_t> nparam, nphi;  // per PR chunk: sigma_t and phi element counts
  std::vector<int32_t> sdoms;         // every subdomain, in SdomId order
  struct Exit {
    int32_t sdom, axis;
    double *fwd_plane, *adj_plane;
    int64_t n;
  };
  std::vector<Exit> exits;            // plane chunks whose faces leave the domain
  double *slab = nullptr;             // one allocation carved into the buffers below
  std::vector<double *> par;          // [field][2][chunk]: sigma_t, s
  std::vector<double *> ckpt;         // [field][checkpoint][chunk]: phi_{j*ck}
  std::vector<double *> blk;          // [ck-1][chunk]: phi of one checkpoint interval, recomputed
  std::vector<double *> s_cur;        // [chunk]: s of the field being run (fixed addresses for the kernels)
  std::vector<double *> gt, gs;       // [chunk]: gradient accumulators
  std::vector<uint64_t> version;      // per field: bumped by kr_diff_set_params (0 = never set)
  std::vector<uint64_t> fwd_version;  // per field: the version its checkpoints were taken with (0 = none)
  uint64_t next_version = 1;
};

namespace {

int dfail(int code, const std::string &msg) {
  set_error(msg);
  return code;
}

#define KR_TRY(expr)          \
  do {                        \
    const int _rc = (expr);   \
    if (_rc != KR_OK) return _rc; \
  } while (0)

double *param(kr_diff *d, int f, int which, int c) { return d->par[((size_t)f * 2 + which) * d->npr + c]; }

hipStream_t stream(kr_diff *d) { return d->fwd->ctx->stream; }

// sigma_t of field f into both problems' sigt_zonal (what the sweeps read),
// s of field f into the kernels' fixed s arrays.
int load_params(kr_diff *d, int f) {
  for (int c = 0; c < d->npr; ++c) {
    KR_TRY(launch_copy(stream(d), (double *)d->fwd->prs[c].sigt_zonal, param(d, f, 0, c), d->nparam[c]));
    KR_TRY(launch_copy(stream(d), (double *)d->adj->prs[c].sigt_zonal, param(d, f, 0, c), d->nparam[c]));
    KR_TRY(launch_copy(stream(d), d->s_cur[c], param(d, f, 1, c), d->nparam[c]));
  }
  return KR_OK;
}

int set_phi(kr_diff *d, const std::vector<double *> &src, size_t first) {
  for (int c = 0; c < d->npr; ++c)
    KR_TRY(launch_copy(stream(d), d->fwd->prs[c].phi, src[first + c], d->nphi[c]));
  return KR_OK;
}

int save_phi(kr_diff *d, std::vector<double *> &dst, size_t first) {
  for (int c = 0; c < d->npr; ++c)
    KR_TRY(launch_copy(stream(d), dst[first + c], d->fwd->prs[c].phi, d->nphi[c]));
  return KR_OK;
}

// From fwd.phi = phi_n: phi_out, rhs and psi_{n+1}; phi stays phi_n.
int tail(kr_diff *d) {
  KR_TRY(kr_scattering_zonal(d->fwd, d->s_cur.data(), 0));
  KR_TRY(kr_source(d->fwd, d->strength));
  KR_TRY(kr_lplustimes(d->fwd, 0));
  return kr_sweep_solver(d->fwd, d->sdoms.data(), d->nsdom);
}

// phi_n -> phi_{n+1} in fwd.phi.
int step(kr_diff *d) {
  KR_TRY(tail(d));
  return kr_ltimes(d->fwd, 0);
}

int check_range(kr_diff *d, int32_t first, int32_t num, const char *who) {
  if (!d) return dfail(KR_ERR_INVALID, std::string(who) + ": null handle");
  if (num < 1 || first < 0 || first > d->F - num)
    return dfail(KR_ERR_INVALID, std::string(who) + ": field range outside [0, num_fields)");
  return KR_OK;
}

}  // namespace

extern "C" {

int kr_diff_create(kr_problem *fwd, kr_problem *adj, int32_t iterations, int32_t checkpoint_every,
                   int32_t num_fields, double source_strength, kr_diff **out) {
  if (!fwd || !adj || !out || fwd == adj) return dfail(KR_ERR_INVALID, "kr_diff_create: bad problem arguments");
  if (iterations < 1 || checkpoint_every < 1 || num_fields < 1)
    return dfail(KR_ERR_INVALID, "kr_diff_create: iterations, checkpoint_every and num_fields must be >= 1");
  if (fwd->ctx != adj->ctx) return dfail(KR_ERR_INVALID, "kr_diff_create: problems on different contexts");
  if (fwd->nesting != adj->nesting || fwd->nsdom != adj->nsdom || fwd->npr != adj->npr || fwd->M != adj->M)
    return dfail(KR_ERR_INVALID, "kr_diff_create: forward and adjoint problems differ in shape");
  for (int s = 0; s < fwd->nsdom; ++s) {
    const DevSdom &x = fwd->h_sd[s], &y = adj->h_sd[s];
    if (x.nd != y.nd || x.ng != y.ng || x.ni != y.ni || x.nj != y.nj || x.nk != y.nk || x.pr != y.pr)
      return dfail(KR_ERR_INVALID, "kr_diff_create: forward and adjoint problems differ in shape");
    if (x.psi == y.psi || x.rhs == y.rhs || x.plane[0] == y.plane[0] || x.plane[1] == y.plane[1] ||
        x.plane[2] == y.plane[2])
      return dfail(KR_ERR_INVALID, "kr_diff_create: the adjoint problem needs its own psi, rhs and planes");
  }
  for (int c = 0; c < fwd->npr; ++c)
    if (fwd->prs[c].phi == adj->prs[c].phi || fwd->prs[c].phi_out == adj->prs[c].phi_out)
      return dfail(KR_ERR_INVALID, "kr_diff_create: the adjoint problem needs its own phi and phi_out");

  kr_diff *d = new kr_diff();
  d->fwd = fwd;
  d->adj = adj;
  d->K = iterations;
  d->ck = checkpoint_every;
  d->nck = (iterations + checkpoint_every - 1) / checkpoint_every;
  d->F = num_fields;
  d->npr = fwd->npr;
  d->nsdom = fwd->nsdom;
  d->strength = source_strength;
  d->version.assign(num_fields, 0);
  d->fwd_version.assign(num_fields, 0);
  for (int s = 0; s < d->nsdom; ++s) d->sdoms.push_back(s);
  for (int c = 0; c < d->npr; ++c) {
    const DevPR &P = fwd->h_pr[c];
    d->nparam.push_back((int64_t)P.ng * P.nz);
    d->nphi.push_back((int64_t)fwd->M * P.ng * P.nz);
  }
  for (int s = 0; s < d->nsdom; ++s) {
    const DevSdom &S = fwd->h_sd[s];
    const int64_t faces[3] = {(int64_t)S.nj * S.nk, (int64_t)S.ni * S.nk, (int64_t)S.ni * S.nj};
    for (int a = 0; a < 3; ++a)
      if (S.down[a] < 0)
        d->exits.push_back({s, a, S.plane[a], adj->h_sd[s].plane[a], (int64_t)S.nd * S.ng * faces[a]});
  }
  // carve: per field 2*npr parameter arrays and nck*npr checkpoints, then
  // (ck-1)*npr interval buffers and 3*npr fixed arrays; 32-double alignment
  auto rnd = [](int64_t n) { return (n + 31) / 32 * 32; };
  int64_t total = 0;
  std::vector<std::pair<std::vector<double *> *, int64_t>> plan;  // (list, offset of the entry)
  auto add = [&](std::vector<double *> &list, int64_t n) {
    plan.push_back({&list, total});
    total += rnd(std::max<int64_t>(n, 1));
  };
  for (int f = 0; f < d->F; ++f)
    for (int w = 0; w < 2; ++w)
      for (int c = 0; c < d->npr; ++c) add(d->par, d->nparam[c]);
  for (int f = 0; f < d->F; ++f)
    for (int j = 0; j < d->nck; ++j)
      for (int c = 0; c < d->npr; ++c) add(d->ckpt, d->nphi[c]);
  for (int j = 0; j + 1 < d->ck; ++j)
    for (int c = 0; c < d->npr; ++c) add(d->blk, d->nphi[c]);
  for (int c = 0; c < d->npr; ++c) add(d->s_cur, d->nparam[c]);
  for (int c = 0; c < d->npr; ++c) add(d->gt, d->nparam[c]);
  for (int c = 0; c < d->npr; ++c) add(d->gs, d->nparam[c]);
  const hipError_t e = hipMalloc((void **)&d->slab, (size_t)total * sizeof(double));
  if (e != hipSuccess) {
    d->slab = nullptr;
    delete d;
    return dfail(KR_ERR_NOMEM, std::string("kr_diff_create: ") + hipGetErrorString(e));
  }
  for (auto &pl : plan) pl.first->push_back(d->slab + pl.second);
  *out = d;
  return KR_OK;
}

int kr_diff_destroy(kr_diff *d) {
  if (!d) return KR_OK;
  // hipFree waits for the work in flight; the problems are not touched (they
  // may already be destroyed)
  if (d->slab) hipFree(d->slab);
  delete d;
  return KR_OK;
}

int kr_diff_num_exits(kr_diff *d, int32_t *num) {
  if (!d || !num) return dfail(KR_ERR_INVALID, "kr_diff_num_exits: null argument");
  *num = (int32_t)d->exits.size();
  return KR_OK;
}

int kr_diff_exit(kr_diff *d, int32_t e, int32_t *sdom, int32_t *axis, int64_t *num_elem) {
  if (!d || e < 0 || e >= (int32_t)d->exits.size()) return dfail(KR_ERR_INVALID, "kr_diff_exit: bad argument");
  if (sdom) *sdom = d->exits[e].sdom;
  if (axis) *axis = d->exits[e].axis;
  if (num_elem) *num_elem = d->exits[e].n;
  return KR_OK;
}

int kr_diff_set_params(kr_diff *d, int32_t field, const double *const *sigt, const double *const *s) {
  KR_TRY(check_range(d, field, 1, "kr_diff_set_params"));
  if (!sigt || !s) return dfail(KR_ERR_INVALID, "kr_diff_set_params: null pointer array");
  for (int c = 0; c < d->npr; ++c)
    if (!sigt[c] || !s[c]) return dfail(KR_ERR_INVALID, "kr_diff_set_params: null chunk pointer");
  for (int c = 0; c < d->npr; ++c) {
    KR_TRY(launch_copy(stream(d), param(d, field, 0, c), sigt[c], d->nparam[c]));
    KR_TRY(launch_copy(stream(d), param(d, field, 1, c), s[c], d->nparam[c]));
  }
  KR_HIP_CHECK(hipStreamSynchronize(stream(d)));  // the caller may reuse its arrays on return
  d->version[field] = d->next_version++;
  return KR_OK;
}

int kr_diff_forward(kr_diff *d, int32_t first, int32_t num, double *const *phi_k, double *const *faces_out) {
  KR_TRY(check_range(d, first, num, "kr_diff_forward"));
  for (int f = first; f < first + num; ++f)
    if (!d->version[f]) return dfail(KR_ERR_INVALID, "kr_diff_forward: parameters of a field were never set");
  const size_t ne = d->exits.size();
  for (int f = first; f < first + num; ++f) {
    d->fwd_version[f] = 0;
    KR_TRY(load_params(d, f));
    for (int c = 0; c < d->npr; ++c) KR_TRY(launch_fill(stream(d), d->fwd->prs[c].phi, d->nphi[c], 0.0));  // L 0
    for (int n = 0; n < d->K; ++n) {
      if (n % d->ck == 0) KR_TRY(save_phi(d, d->ckpt, ((size_t)f * d->nck + n / d->ck) * d->npr));
      KR_TRY(step(d));
    }
    d->fwd_version[f] = d->version[f];
    const size_t fl = (size_t)(f - first);
    if (phi_k)
      for (int c = 0; c < d->npr; ++c)
        if (phi_k[fl * d->npr + c])
          KR_TRY(launch_copy(stream(d), phi_k[fl * d->npr + c], d->fwd->prs[c].phi, d->nphi[c]));
    if (faces_out)
      for (size_t e = 0; e < ne; ++e)
        if (faces_out[fl * ne + e])
          KR_TRY(launch_copy(stream(d), faces_out[fl * ne + e], d->exits[e].fwd_plane, d->exits[e].n));
  }
  return KR_OK;
}

int kr_diff_backward(kr_diff *d, int32_t first, int32_t num, const double *const *phibar,
                     const double *const *facebar, double *const *grad_sigt, double *const *grad_s) {
  KR_TRY(check_range(d, first, num, "kr_diff_backward"));
  for (int f = first; f < first + num; ++f)
    if (!d->fwd_version[f] || d->fwd_version[f] != d->version[f])
      return dfail(KR_ERR_INVALID, "kr_diff_backward: a field has no forward pass of its current parameters");
  const size_t ne = d->exits.size();
  for (int f = first; f < first + num; ++f) {
    const size_t fl = (size_t)(f - first);
    KR_TRY(load_params(d, f));
    for (int c = 0; c < d->npr; ++c) {
      const double *pb = phibar ? phibar[fl * d->npr + c] : nullptr;
      if (pb)
        KR_TRY(launch_copy(stream(d), d->adj->prs[c].phi, pb, d->nphi[c]));
      else
        KR_TRY(launch_fill(stream(d), d->adj->prs[c].phi, d->nphi[c], 0.0));
      KR_TRY(launch_fill(stream(d), d->gt[c], d->nparam[c], 0.0));
      KR_TRY(launch_fill(stream(d), d->gs[c], d->nparam[c], 0.0));
    }
    KR_TRY(kr_ltimes_adjoint(d->adj));  // psibar_K
    for (int n = d->K - 1; n >= 0; --n) {
      const int b = n / d->ck, c0 = b * d->ck;
      const size_t ck_first = ((size_t)f * d->nck + b) * d->npr;
      if (n == std::min(d->K, c0 + d->ck) - 1) {
        // last iteration of an interval: phi_{c0+1} .. phi_n once from the checkpoint
        KR_TRY(set_phi(d, d->ckpt, ck_first));
        for (int j = c0; j < n; ++j) {
          KR_TRY(step(d));
          KR_TRY(save_phi(d, d->blk, (size_t)(j - c0) * d->npr));
        }
      } else if (n == c0) {
        KR_TRY(set_phi(d, d->ckpt, ck_first));
      } else {
        KR_TRY(set_phi(d, d->blk, (size_t)(n - c0 - 1) * d->npr));
      }
      KR_TRY(tail(d));  // psi_{n+1}; fwd.phi still phi_n
      if (n == d->K - 1 && facebar) {
        for (size_t e = 0; e < ne; ++e) {  // adjoint inflow: the exiting faces' cotangent
          const double *fb = facebar[fl * ne + e];
          if (fb)
            KR_TRY(launch_copy(stream(d), d->exits[e].adj_plane, fb, d->exits[e].n));
          else
            KR_TRY(launch_fill(stream(d), d->exits[e].adj_plane, d->exits[e].n, 0.0));
        }
        KR_TRY(kr_sweep_solver_adjoint_seeded(d->adj));  // rhsbar_{K-1}
      } else {
        KR_TRY(kr_sweep_solver_adjoint(d->adj));  // rhsbar_n
      }
      KR_TRY(kr_grad_sigt(d->fwd, d->adj, d->gt.data(), 1));
      KR_TRY(kr_lplustimes_adjoint(d->adj));  // phi_outbar_n
      KR_TRY(kr_grad_sigs_zonal(d->fwd, d->adj, d->gs.data(), 1));
      KR_TRY(kr_scattering_zonal_adjoint(d->adj, d->s_cur.data(), 0));  // phibar_n
      KR_TRY(kr_ltimes_adjoint(d->adj));  // psibar_n
    }
    for (int c = 0; c < d->npr; ++c) {
      if (grad_sigt && grad_sigt[fl * d->npr + c])
        KR_TRY(launch_copy(stream(d), grad_sigt[fl * d->npr + c], d->gt[c], d->nparam[c]));
      if (grad_s && grad_s[fl * d->npr + c])
        KR_TRY(launch_copy(stream(d), grad_s[fl * d->npr + c], d->gs[c], d->nparam[c]));
    }
  }
  return KR_OK;
}

}  // extern "C"
