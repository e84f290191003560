// C-ABI over the host mirror (kr_host_* in include/kripke_amd.h).
// C++ exceptions stop here and become status codes.
#include <chrono>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "kripke_host.h"

using namespace Kripke;
using namespace Kripke::Core;

namespace kr {
void set_error(const std::string &msg);
}

struct kr_host_problem {
  DataStore ds;
  kr_context *ctx = nullptr;
  int32_t nsdom = 0, npr = 0, nr = 0;
  bool block_jacobi = false;  // InputVariables::parallel_method == PMETHOD_BJ
};

namespace {

template <class F>
int guarded(F &&f) {
  try {
    return f();
  } catch (const KripkeError &e) {
    kr::set_error(e.what());
    return e.code;
  } catch (const std::domain_error &e) {
    kr::set_error(e.what());
    return KR_ERR_INVALID;
  } catch (const std::exception &e) {
    kr::set_error(e.what());
    return KR_ERR_INVALID;
  }
}

// Field chunk -> doubles, from HBM when the field lives there.
template <class T>
int get_chunk(kr_context *ctx, Field<T> &f, int32_t chunk, double *dst, int64_t *n) {
  if (chunk < 0 || (size_t)chunk >= f.numChunks()) {
    kr::set_error("chunk out of range");
    return KR_ERR_INVALID;
  }
  const SdomId s = f.chunkSdom((size_t)chunk);
  const size_t cnt = f.size(s);
  if (n) *n = (int64_t)cnt;
  if (!dst) return KR_OK;
  std::vector<T> tmp(cnt);
  if (f.onDevice()) {
    int rc = kr_memcpy_dtoh(ctx, tmp.data(), f.device(s), cnt * sizeof(T));
    if (rc) return rc;
  } else if (f.onHost()) {
    memcpy(tmp.data(), f.host(s), cnt * sizeof(T));
  } else {
    kr::set_error("field has no storage in this problem (host-only problem?)");
    return KR_ERR_INVALID;
  }
  for (size_t i = 0; i < cnt; ++i) dst[i] = (double)tmp[i];
  return KR_OK;
}

// Nesting choices of kr_host_autotune_nesting, per context and problem shape.
std::mutex g_nest_mu;
std::map<std::pair<const kr_context *, std::vector<int64_t>>, std::pair<int32_t, double>> g_nest_cache;

}  // namespace

extern "C" {

int kr_host_create(kr_context *ctx, const kr_input_vars *in, kr_host_problem **out) {
  if (!in || !out) {
    kr::set_error("kr_host_create: null argument");
    return KR_ERR_INVALID;
  }
  kr_host_problem *hp = new kr_host_problem();
  hp->ctx = ctx;
  hp->block_jacobi = in->parallel_method == 1;
  int rc = guarded([&] {
    generateProblem(hp->ds, *in, ctx);
    auto &ps = hp->ds.getVariable<PartitionSpace>("pspace");
    hp->nsdom = (int32_t)ps.getNumSubdomains(SPACE_PQR);
    hp->npr = (int32_t)ps.getNumSubdomains(SPACE_PR);
    hp->nr = (int32_t)ps.getNumSubdomains(SPACE_R);
    return KR_OK;
  });
  if (rc != KR_OK) {
    delete hp;
    return rc;
  }
  *out = hp;
  return KR_OK;
}

int kr_host_destroy(kr_host_problem *hp) {
  if (hp && hp->ctx) kr_synchronize(hp->ctx);
  delete hp;
  return KR_OK;
}

int kr_host_num_sdoms(kr_host_problem *hp, int32_t *ns, int32_t *npr, int32_t *nr) {
  if (!hp) return KR_ERR_INVALID;
  if (ns) *ns = hp->nsdom;
  if (npr) *npr = hp->npr;
  if (nr) *nr = hp->nr;
  return KR_OK;
}

kr_problem *kr_host_kernel_problem(kr_host_problem *hp) {
  if (!hp || !hp->ds.isVariableDefined("kr_problem")) return nullptr;
  return hp->ds.getVariable<KernelProblem>("kr_problem").prob;
}

int kr_host_field_get(kr_host_problem *hp, const char *name, int32_t chunk, double *dst, int64_t *n) {
  if (!hp || !name) return KR_ERR_INVALID;
  return guarded([&] {
    const std::string nm(name);
    if (hp->ds.isVariableType<Field<double>>(nm))
      return get_chunk(hp->ctx, hp->ds.getVariable<Field<double>>(nm), chunk, dst, n);
    if (hp->ds.isVariableType<Field<int32_t>>(nm))
      return get_chunk(hp->ctx, hp->ds.getVariable<Field<int32_t>>(nm), chunk, dst, n);
    if (hp->ds.isVariableType<Field<int64_t>>(nm))
      return get_chunk(hp->ctx, hp->ds.getVariable<Field<int64_t>>(nm), chunk, dst, n);
    kr::set_error("no field named '" + nm + "'");
    return KR_ERR_INVALID;
  });
}

int kr_host_field_set(kr_host_problem *hp, const char *name, int32_t chunk, const double *src) {
  if (!hp || !name || !src) return KR_ERR_INVALID;
  return guarded([&] {
    auto &f = hp->ds.getVariable<Field<double>>(name);
    if (chunk < 0 || (size_t)chunk >= f.numChunks()) {
      kr::set_error("chunk out of range");
      return KR_ERR_INVALID;
    }
    const SdomId s = f.chunkSdom((size_t)chunk);
    if (f.onHost()) memcpy(f.host(s), src, f.size(s) * sizeof(double));
    if (f.onDevice()) return kr_memcpy_htod(hp->ctx, f.device(s), src, f.size(s) * sizeof(double));
    return KR_OK;
  });
}

int kr_host_field_device(kr_host_problem *hp, const char *name, int32_t chunk, void **ptr, int64_t *n) {
  if (!hp || !name || !ptr) return KR_ERR_INVALID;
  return guarded([&] {
    auto &f = hp->ds.getVariable<Field<double>>(name);
    if (chunk < 0 || (size_t)chunk >= f.numChunks() || !f.onDevice()) {
      kr::set_error("chunk out of range or field not on the device");
      return KR_ERR_INVALID;
    }
    const SdomId s = f.chunkSdom((size_t)chunk);
    *ptr = f.device(s);
    if (n) *n = (int64_t)f.size(s);
    return KR_OK;
  });
}

int kr_host_init_psi(kr_host_problem *hp) {
  if (!hp || !hp->ctx) {
    kr::set_error("kr_host_init_psi: problem has no device");
    return KR_ERR_INVALID;
  }
  return guarded([&] {
    Kernel::kConst(hp->ds.getVariable<Field<double>>("psi"), 0.0);
    return KR_OK;
  });
}

int kr_host_iterate(kr_host_problem *hp, int32_t mode, double *population) {
  if (!hp || !hp->ctx) {
    kr::set_error("kr_host_iterate: problem has no device");
    return KR_ERR_INVALID;
  }
  return guarded([&] {
    const double p = SteadyStateIteration(hp->ds, mode, hp->block_jacobi);
    if (population) *population = p;
    return KR_OK;
  });
}

int kr_host_solve(kr_host_problem *hp, int32_t niter, int32_t mode, double *pops) {
  int rc = kr_host_init_psi(hp);
  if (rc) return rc;
  for (int32_t it = 0; it < niter; ++it) {
    rc = kr_host_iterate(hp, mode, pops ? pops + it : nullptr);
    if (rc) return rc;
  }
  return KR_OK;
}

// Times one source iteration of the problem in each of the six nestings
// (each with its sweep shape tuned) and returns the fastest.  The choice is
// cached per context for every problem of the same extents and sets.
int kr_host_autotune_nesting(kr_context *ctx, const kr_input_vars *in, int32_t *nesting, double *ms) {
  if (!ctx || !in || !nesting) {
    kr::set_error("kr_host_autotune_nesting: null argument");
    return KR_ERR_INVALID;
  }
  const std::vector<int64_t> key = {in->nx, in->ny, in->nz, in->num_directions, in->quad_num_polar,
                                    in->quad_num_azimuthal, in->num_groups, in->legendre_order, in->num_dirsets,
                                    in->num_groupsets, in->num_zonesets[0], in->num_zonesets[1], in->num_zonesets[2],
                                    in->parallel_method};
  {
    std::lock_guard<std::mutex> lk(g_nest_mu);
    auto it = g_nest_cache.find({ctx, key});
    if (it != g_nest_cache.end()) {
      *nesting = it->second.first;
      if (ms) *ms = it->second.second;
      return KR_OK;
    }
  }
  int32_t best = -1;
  double best_ms = 1e300;
  std::string last_error;
  for (int32_t lay = KR_DGZ; lay <= KR_ZGD; ++lay) {
    kr_input_vars v = *in;
    v.layout = lay;
    kr_host_problem *hp = nullptr;
    int rc = kr_host_create(ctx, &v, &hp);
    if (rc == KR_OK) rc = kr_problem_autotune(kr_host_kernel_problem(hp), nullptr);
    if (rc == KR_OK) rc = kr_host_init_psi(hp);
    if (rc == KR_OK) rc = kr_host_iterate(hp, 1, nullptr);  // warm-up
    double t = 1e300;
    for (int rep = 0; rep < 2 && rc == KR_OK; ++rep) {
      const auto t0 = std::chrono::steady_clock::now();
      rc = kr_host_iterate(hp, 1, nullptr);  // ends with the population read-back
      t = std::min(t, std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
    }
    if (rc != KR_OK) last_error = kr_last_error();
    if (hp) kr_host_destroy(hp);
    if (rc == KR_ERR_UNSUPPORTED) continue;  // a nesting whose shapes the kernels do not cover
    if (rc != KR_OK) {
      kr::set_error("kr_host_autotune_nesting: " + last_error);
      return rc;
    }
    if (t < best_ms) {
      best_ms = t;
      best = lay;
    }
  }
  if (best < 0) {
    kr::set_error("kr_host_autotune_nesting: no nesting supports this problem (" + last_error + ")");
    return KR_ERR_UNSUPPORTED;
  }
  {
    std::lock_guard<std::mutex> lk(g_nest_mu);
    g_nest_cache[{ctx, key}] = {best, best_ms};
  }
  *nesting = best;
  if (ms) *ms = best_ms;
  return KR_OK;
}

}  // extern "C"
