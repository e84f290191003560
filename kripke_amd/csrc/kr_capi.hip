// C-ABI layer: contexts, device memory, problem registration and the
// Kernel:: entry points declared in include/kripke_amd.h.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "kr_state.h"

namespace kr {
static thread_local std::string g_last_error;
void set_error(const std::string &msg) { g_last_error = msg; }
}  // namespace kr

using namespace kr;

// Scratch of kr_sweep_solver_population_noalloc, the entry point that
// allocates nothing: the fused population's per-work-unit partials followed
// by their sum.  Every device holds its own copy for the life of the process;
// host code reaches it only through the device address hipGetSymbolAddress
// returns on that device (kr_context::noalloc_scratch), never through the
// array's name.
static __device__ __align__(256) double scratch[KR_NOALLOC_SCRATCH_DOUBLES];

// Per-device addresses of `scratch`, looked up once per device at context
// creation with that device current.
static std::mutex g_scratch_mu;
static std::map<int, double *> g_scratch_addr;

static int noalloc_scratch_address(int device, double **out) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  auto it = g_scratch_addr.find(device);
  if (it == g_scratch_addr.end()) {
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
      return KR_ERR_HIP;
    }
    it = g_scratch_addr.emplace(device, (double *)ptr).first;
  }
  *out = it->second;
  return KR_OK;
}

// A batch whose member problem is destroyed keeps its handle valid (so that
// kr_batch_destroy still works) but refuses kr_batch_iterate.
static void batch_forget_problem(kr_batch *b, kr_problem *p);

namespace {

int fail(int code, const std::string &msg) {
  set_error(msg);
  return code;
}

template <class T>
int upload(kr_context *ctx, T **dptr, const std::vector<T> &v) {
  if (v.empty()) {
    *dptr = nullptr;
    return KR_OK;
  }
  KR_HIP_CHECK(hipMalloc((void **)dptr, v.size() * sizeof(T)));
  KR_HIP_CHECK(hipMemcpyAsync(*dptr, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, ctx->stream));
  KR_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  return KR_OK;
}

template <class T>
int download(kr_context *ctx, std::vector<T> &v, const T *dptr, size_t n) {
  v.resize(n);
  if (n == 0) return KR_OK;
  KR_HIP_CHECK(hipMemcpyAsync(v.data(), dptr, n * sizeof(T), hipMemcpyDeviceToHost, ctx->stream));
  KR_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  return KR_OK;
}

LaunchCtx launch_ctx(kr_problem *p) {
  LaunchCtx L{};
  L.stream = p->ctx->stream;
  L.sd = p->d_sd;
  L.pr = p->d_pr;
  L.r = p->d_r;
  L.pr_sdoms = p->d_pr_sdoms;
  L.sigs = p->sigs;
  L.m2l = p->m2l;
  for (int i = 0; i < 4; ++i) L.sigs_str[i] = p->sigs_str[i];
  L.sigs_dims[0] = p->nmat;
  L.sigs_dims[1] = p->L1;
  L.sigs_dims[2] = p->G;
  L.sigs_dims[3] = p->G;
  L.sigs_flags = p->d_sigs_flags;
  L.sigs_t = p->d_sigs_t;
  L.num_pr = p->npr;
  L.num_moments = p->M;
  L.num_r = p->nr;
  L.max_gz = p->max_gz;
  L.max_gmix = p->max_gmix;
  L.max_g = p->max_g;
  L.max_z = p->max_z;
  L.nd = p->h_sd.empty() ? 1 : p->h_sd[0].nd;
  L.dfast = (p->nesting == KR_GZD || p->nesting == KR_ZGD) ? 1 : 0;
  L.gfast = (p->nesting == KR_DZG || p->nesting == KR_ZDG) ? 1 : 0;
  L.lds_max = p->ctx->lds_max;
  return L;
}

// Scoped event timing of one entry point on the context stream.
struct CallTimer {
  kr_problem *p;
  TimedCall tc;
  bool on;
  CallTimer(kr_problem *pp, const char *name) : p(pp), on(pp->timing != 0) {
    if (!on) return;
    tc.name = name;
    hipEventCreate(&tc.start);
    hipEventCreate(&tc.stop);
    hipEventRecord(tc.start, p->ctx->stream);
  }
  ~CallTimer() {
    if (!on) return;
    hipEventRecord(tc.stop, p->ctx->stream);
    p->pending.push_back(tc);
  }
};

int flush_timers(kr_problem *p) {
  if (p->pending.empty()) return KR_OK;
  KR_HIP_CHECK(hipStreamSynchronize(p->ctx->stream));
  for (auto &tc : p->pending) {
    float ms = 0.f;
    hipEventElapsedTime(&ms, tc.start, tc.stop);
    auto &t = p->timers[tc.name];
    t.first += ms * 1e-3;
    t.second += 1;
    hipEventDestroy(tc.start);
    hipEventDestroy(tc.stop);
  }
  p->pending.clear();
  return KR_OK;
}

// Units for the pair groups of one subdomain, R pair groups each.
void append_units(std::vector<SweepUnit> &u, int sdom, int32_t pgs, int32_t R) {
  for (int32_t b = 0; b < pgs; b += R) u.push_back(SweepUnit{sdom, b, std::min(R, pgs - b), 0});
}

// Dispatch order of one launch's units [begin, begin + n).  Units that read
// the same sigma_t rows in the same order (same sigma_t chunk, same group
// sequence: with the group index fastest it is fixed by the first pair
// modulo the group count, otherwise by the first pair) are placed back to
// back on one XCD: workgroups go to the 8 XCDs round-robin, so slot j*8 + x
// runs on XCD x, and sorted unit i is given XCD i / ceil(n/8) and the next
// free slot there.  Such units start together and stream the same rows
// through that XCD's L2.  Units of a launch are independent, so only speed
// depends on this order.
void order_units_xcd(const kr_problem *p, const SweepShape &sh, std::vector<SweepUnit> &u, size_t begin, size_t n) {
  if (n < 2) return;
  struct Key {
    uintptr_t sig;
    int64_t gseq;
    int32_t cnt;
    size_t idx;
  };
  std::vector<Key> k(n);
  for (size_t i = 0; i < n; ++i) {
    const SweepUnit &U = u[begin + i];
    const DevSdom &S = p->h_sd[U.sdom];
    const int64_t p0 = (int64_t)U.pg_begin * sh.Q;
    k[i] = Key{(uintptr_t)S.sig, sh.inner_is_g ? p0 % S.ng : p0, U.pg_count, i};
  }
  std::sort(k.begin(), k.end(), [](const Key &a, const Key &b) {
    if (a.sig != b.sig) return a.sig < b.sig;
    if (a.gseq != b.gseq) return a.gseq < b.gseq;
    if (a.cnt != b.cnt) return a.cnt < b.cnt;
    return a.idx < b.idx;
  });
  const size_t q = n / 8, rem = n % 8;
  std::vector<SweepUnit> out(n);
  size_t i = 0;
  for (size_t x = 0; x < 8; ++x) {
    const size_t cx = q + (x < rem ? 1 : 0);
    for (size_t j = 0; j < cx; ++j) out[j * 8 + x] = u[begin + k[i++].idx];
  }
  std::copy(out.begin(), out.end(), u.begin() + begin);
}

// The sweep kernels address each unit's slice of psi/rhs, the planes and
// sigma_t through buffer resources with 32-bit byte offsets: every unit's span
// (from its first pair's first element to its last pair's last element) must
// stay below 2 GiB.  Offsets grow with the pair index in every nesting.
int check_unit_spans(const kr_problem *p, const std::vector<SweepUnit> &units, const SweepShape &sh) {
  const int64_t lim = 0x7FFFFFF8LL / 8;
  for (const SweepUnit &u : units) {
    const DevSdom &S = p->h_sd[u.sdom];
    const int64_t first = (int64_t)u.pg_begin * sh.Q, last = (int64_t)(u.pg_begin + u.pg_count) * sh.Q - 1;
    auto off = [&](int64_t pidx, int64_t sa, int64_t sg) {
      const int64_t d = sh.inner_is_g ? pidx / S.ng : pidx % S.nd;
      const int64_t g = sh.inner_is_g ? pidx % S.ng : pidx / S.nd;
      return d * sa + g * sg;
    };
    if (off(last, S.fa, S.fg) - off(first, S.fa, S.fg) + (S.nz - 1) * S.fz + 1 > lim)
      return fail(KR_ERR_UNSUPPORTED, "sweep: a work unit's psi slice exceeds 2 GiB");
    const int64_t nf[3] = {(int64_t)S.nj * S.nk, (int64_t)S.ni * S.nk, (int64_t)S.ni * S.nj};
    for (int x = 0; x < 3; ++x)
      if (off(last, S.pa[x], S.pg[x]) - off(first, S.pa[x], S.pg[x]) + (nf[x] - 1) * S.pf[x] + 1 > lim)
        return fail(KR_ERR_UNSUPPORTED, "sweep: a work unit's plane slice exceeds 2 GiB");
    if ((int64_t)S.ng * S.nz > lim) return fail(KR_ERR_UNSUPPORTED, "sweep: sigma_t chunk exceeds 2 GiB");
    const int64_t strides[] = {S.fa, S.fg, S.fz, S.sg, S.sz, S.pa[0], S.pa[1], S.pa[2], S.pg[0], S.pg[1], S.pg[2]};
    for (int64_t st : strides)
      if (st >= (1LL << 31)) return fail(KR_ERR_UNSUPPORTED, "sweep: a field stride exceeds 2^31 elements");
  }
  return KR_OK;
}

// Items per work unit for a level of total_pgs items.  The row-scan kernel
// takes one unit per resident workgroup slot.  The systolic kernel takes
// eight: its per-unit times vary, and a level of many short units balances
// across the CUs and XCDs.  On 64^3 zones x 128 groups x 128 directions in
// 2x2x2 zone sets (R = 12 -> 1 on the middle levels) its sweep took
// 32.8 -> 22.9 ms (DZG) and 30.7 -> 21.8 ms (GZD) in the autotuner's
// sweep-only timings, while R other than one unit per slot made the
// row-scan sweep slower.
int32_t pick_R(int64_t total_pgs, const SweepShape &sh) {
  const int nt = sh.kind == 1 ? sh.RG * sh.PK : sh.Q * sh.PJ * sh.PK;
  const int64_t slots = 256LL * std::max(1, 1024 / std::max(nt, 1));  // resident workgroups, one per 1024 threads
  const int64_t per_slot = sh.kind == 1 ? 1 : 8;
  int64_t R = total_pgs / std::max<int64_t>(slots * per_slot, 1);
  if (R < 1) R = 1;
  if (R > 64) R = 64;
  return (int32_t)R;
}

// Dependency levels of a subdomain list from the on-rank adjacency.
int sweep_levels(const kr_problem *p, const std::vector<int32_t> &list, std::vector<std::vector<int32_t>> &levels) {
  std::vector<int> level(p->nsdom, -1), inlist(p->nsdom, 0);
  for (int32_t s : list) {
    if (s < 0 || s >= p->nsdom) return fail(KR_ERR_INVALID, "sweep: subdomain id out of range");
    inlist[s] = 1;
  }
  for (int32_t s : list)
    for (int dim = 0; dim < 3; ++dim) {
      int u = p->sdoms[s].upwind[dim];
      if (u >= 0 && !inlist[u]) return fail(KR_ERR_INVALID, "sweep: an upwind subdomain is not in the list");
    }
  bool changed = true;
  int guard = 0;
  while (changed) {
    changed = false;
    if (++guard > p->nsdom + 2) return fail(KR_ERR_INVALID, "sweep: dependency cycle");
    for (int32_t s : list) {
      int lv = 0;
      bool ready = true;
      for (int dim = 0; dim < 3; ++dim) {
        int u = p->sdoms[s].upwind[dim];
        if (u < 0) continue;
        if (level[u] < 0) { ready = false; break; }
        lv = std::max(lv, level[u] + 1);
      }
      if (ready && level[s] != lv) {
        level[s] = lv;
        changed = true;
      }
    }
  }
  int maxl = -1;
  for (int32_t s : list) {
    if (level[s] < 0) return fail(KR_ERR_INVALID, "sweep: unresolved dependency");
    maxl = std::max(maxl, level[s]);
  }
  levels.assign(maxl + 1, {});
  for (int32_t s : list) levels[level[s]].push_back(s);
  return KR_OK;
}

}  // namespace

extern "C" {

const char *kr_last_error(void) { return g_last_error.c_str(); }

int kr_device_count(int32_t *count) {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess) n = 0;
  *count = n;
  return KR_OK;
}

int kr_context_create(int32_t device, kr_context **out) {
  if (!out) return fail(KR_ERR_INVALID, "kr_context_create: null out");
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return fail(KR_ERR_NO_DEVICE, "no HIP device available");
  if (device < 0 || device >= n) return fail(KR_ERR_INVALID, "kr_context_create: bad device index");
  KR_HIP_CHECK(hipSetDevice(device));
  kr_context *c = new kr_context();
  c->device = device;
  int lds = 0;
  if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) == hipSuccess && lds > 0)
    c->lds_max = (size_t)lds;
  {
    const int rc = noalloc_scratch_address(device, &c->noalloc_scratch);
    if (rc) {
      delete c;
      return rc;
    }
  }
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete c;
    return fail(KR_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
  }
  *out = c;
  return KR_OK;
}

int kr_context_destroy(kr_context *ctx) {
  if (!ctx) return KR_OK;
  hipStreamSynchronize(ctx->stream);
  hipStreamDestroy(ctx->stream);
  delete ctx;
  return KR_OK;
}

int kr_malloc(kr_context *ctx, size_t bytes, void **dptr) {
  if (!ctx || !dptr) return fail(KR_ERR_INVALID, "kr_malloc: null argument");
  KR_HIP_CHECK(hipSetDevice(ctx->device));
  hipError_t e = hipMalloc(dptr, bytes ? bytes : 8);
  if (e != hipSuccess) return fail(KR_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  return KR_OK;
}

int kr_free(kr_context *ctx, void *dptr) {
  if (!ctx) return fail(KR_ERR_INVALID, "kr_free: null context");
  if (!dptr) return KR_OK;
  KR_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  KR_HIP_CHECK(hipFree(dptr));
  return KR_OK;
}

int kr_memcpy_htod(kr_context *ctx, void *dst, const void *src, size_t bytes) {
  if (!ctx) return fail(KR_ERR_INVALID, "kr_memcpy_htod: null context");
  if (!bytes) return KR_OK;
  KR_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  KR_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  return KR_OK;
}

int kr_memcpy_dtoh(kr_context *ctx, void *dst, const void *src, size_t bytes) {
  if (!ctx) return fail(KR_ERR_INVALID, "kr_memcpy_dtoh: null context");
  if (!bytes) return KR_OK;
  KR_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  KR_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  return KR_OK;
}

int kr_synchronize(kr_context *ctx) {
  if (!ctx) return fail(KR_ERR_INVALID, "kr_synchronize: null context");
  KR_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  return KR_OK;
}

int kr_problem_create(kr_context *ctx, const kr_problem_desc *desc, kr_problem **out) {
  if (!ctx || !desc || !out) return fail(KR_ERR_INVALID, "kr_problem_create: null argument");
  const kr_problem_desc &D = *desc;
  if (D.nesting < 0 || D.nesting > 5) return fail(KR_ERR_INVALID, "kr_problem_create: nesting out of range");
  if (D.num_sdoms <= 0 || D.num_pr <= 0 || D.num_r <= 0 || !D.sdoms || !D.pr || !D.r)
    return fail(KR_ERR_INVALID, "kr_problem_create: empty subdomain tables");
  if (D.num_moments <= 0 || D.num_legendre <= 0 || D.num_moments != D.num_legendre * D.num_legendre)
    return fail(KR_ERR_INVALID, "kr_problem_create: moments must be (L+1)^2");
  if (!D.sigs || !D.moment_to_legendre || D.num_materials <= 0 || D.global_num_groups <= 0)
    return fail(KR_ERR_INVALID, "kr_problem_create: missing scattering data");
  KR_HIP_CHECK(hipSetDevice(ctx->device));

  kr_problem *p = new kr_problem();
  p->ctx = ctx;
  p->nesting = D.nesting;
  p->nsdom = D.num_sdoms;
  p->npr = D.num_pr;
  p->nr = D.num_r;
  p->M = D.num_moments;
  p->L1 = D.num_legendre;
  p->nmat = D.num_materials;
  p->G = D.global_num_groups;
  p->sdoms.assign(D.sdoms, D.sdoms + D.num_sdoms);
  p->prs.assign(D.pr, D.pr + D.num_pr);
  p->rs.assign(D.r, D.r + D.num_r);
  p->sigs = D.sigs;
  p->m2l = D.moment_to_legendre;
  auto bad = [&](const std::string &m) {
    kr_problem_destroy(p);
    return fail(KR_ERR_INVALID, "kr_problem_create: " + m);
  };
  if (p->nr > 0 && hipMalloc((void **)&p->d_vol_info, sizeof(double) * 2 * p->nr) != hipSuccess) {
    p->d_vol_info = nullptr;
    kr_problem_destroy(p);
    return fail(KR_ERR_NOMEM, "kr_problem_create: volume table");
  }

  // --- validate subdomains ------------------------------------------------
  const kr_sdom_desc &s0 = p->sdoms[0];
  std::vector<int> pr_rep(p->npr, -1);
  for (int s = 0; s < p->nsdom; ++s) {
    const kr_sdom_desc &S = p->sdoms[s];
    if (S.num_directions <= 0 || S.num_groups <= 0 || S.ni <= 0 || S.nj <= 0 || S.nk <= 0)
      return bad("subdomain with an empty set");
    if (S.num_directions != s0.num_directions || S.num_groups != s0.num_groups || S.ni != s0.ni ||
        S.nj != s0.nj || S.nk != s0.nk)
      return bad("subdomains must share one shape");
    if (S.pr_chunk < 0 || S.pr_chunk >= p->npr || S.r_chunk < 0 || S.r_chunk >= p->nr)
      return bad("chunk id out of range");
    if (!S.psi || !S.rhs || !S.i_plane || !S.j_plane || !S.k_plane || !S.xcos || !S.ycos || !S.zcos || !S.w ||
        !S.id || !S.jd || !S.kd || !S.ell || !S.ell_plus || !S.dx || !S.dy || !S.dz)
      return bad("missing field pointer");
    for (int dim = 0; dim < 3; ++dim) {
      int u = S.upwind[dim], w = S.downwind[dim];
      if (u >= p->nsdom || w >= p->nsdom || u < -1 || w < -1) return bad("adjacency out of range");
      if (u >= 0 && p->sdoms[u].downwind[dim] != s) return bad("upwind/downwind tables disagree");
    }
    if ((int64_t)S.num_directions * S.num_groups * S.ni * S.nj * S.nk > (int64_t)INT32_MAX)
      return bad("subdomain chunk exceeds 2^31 elements");
    if (pr_rep[S.pr_chunk] < 0) pr_rep[S.pr_chunk] = s;
  }
  for (int c = 0; c < p->npr; ++c) {
    if (pr_rep[c] < 0) return bad("PR chunk without a subdomain");
    if (!p->prs[c].phi || !p->prs[c].phi_out || !p->prs[c].sigt_zonal) return bad("missing PR field pointer");
  }
  for (int c = 0; c < p->nr; ++c) {
    const kr_r_desc &R = p->rs[c];
    if (!R.volume || !R.zone_to_mixelem || !R.zone_to_num_mixelem || !R.mixelem_to_zone ||
        !R.mixelem_to_material || !R.mixelem_to_fraction || R.num_mixelem <= 0)
      return bad("missing zone-set field pointer");
  }

  // --- host copies of small tables (quadrature, mesh, sigs) --------------
  const int nd = s0.num_directions, ng = s0.num_groups, ni = s0.ni, nj = s0.nj, nk = s0.nk;
  const int64_t nz = (int64_t)ni * nj * nk;
  // 2*cos/delta tables, refreshed on device before each sweep, each followed
  // by its uniform-rows flag (k_sweep_coef)
  const int64_t coef_per = (int64_t)nd * (ni + nj + nk) + 1;
  std::vector<double> coef(coef_per * p->nsdom, 0.0);
  p->h_sd.resize(p->nsdom);
  int rc = KR_OK;
  for (int s = 0; s < p->nsdom; ++s) {
    const kr_sdom_desc &S = p->sdoms[s];
    int32_t sgn[3];
    std::vector<int32_t> tmp;
    if ((rc = download(ctx, tmp, S.id, 1))) { kr_problem_destroy(p); return rc; }
    sgn[0] = tmp[0];
    if ((rc = download(ctx, tmp, S.jd, 1))) { kr_problem_destroy(p); return rc; }
    sgn[1] = tmp[0];
    if ((rc = download(ctx, tmp, S.kd, 1))) { kr_problem_destroy(p); return rc; }
    sgn[2] = tmp[0];

    DevSdom &H = p->h_sd[s];
    memset(&H, 0, sizeof(H));
    H.psi = S.psi;
    H.rhs = S.rhs;
    H.plane[0] = S.i_plane;
    H.plane[1] = S.j_plane;
    H.plane[2] = S.k_plane;
    H.sig = p->prs[S.pr_chunk].sigt_zonal;
    H.cosines[0] = S.xcos;
    H.cosines[1] = S.ycos;
    H.cosines[2] = S.zcos;
    H.delta[0] = S.dx;
    H.delta[1] = S.dy;
    H.delta[2] = S.dz;
    H.w = S.w;
    H.ell = S.ell;
    H.ell_plus = S.ell_plus;
    H.volume = p->rs[S.r_chunk].volume;
    H.vol_info = p->d_vol_info + 2 * S.r_chunk;
    H.nd = nd; H.ng = ng; H.ni = ni; H.nj = nj; H.nk = nk; H.nz = (int32_t)nz;
    H.id = sgn[0]; H.jd = sgn[1]; H.kd = sgn[2];
    H.pr = S.pr_chunk; H.r = S.r_chunk; H.gset = S.gset; H.dset = S.dset; H.glo = S.group_lower;
    for (int dim = 0; dim < 3; ++dim) { H.up[dim] = S.upwind[dim]; H.down[dim] = S.downwind[dim]; }
    Str3 f = flux_strides(p->nesting, nd, ng, nz);
    H.fa = f.a; H.fg = f.g; H.fz = f.z;
    sigt_strides(p->nesting, ng, nz, &H.sg, &H.sz);
    const int64_t faces[3] = {(int64_t)nj * nk, (int64_t)ni * nk, (int64_t)ni * nj};
    for (int dim = 0; dim < 3; ++dim) {
      Str3 ps = flux_strides(p->nesting, nd, ng, faces[dim]);
      H.pa[dim] = ps.a; H.pg[dim] = ps.g; H.pf[dim] = ps.z;
    }
  }
  rc = upload(ctx, &p->d_coef, coef);
  if (rc) { kr_problem_destroy(p); return rc; }
  for (int s = 0; s < p->nsdom; ++s) {
    double *A = p->d_coef + coef_per * s;
    p->h_sd[s].coef[0] = A;
    p->h_sd[s].coef[1] = A + (int64_t)nd * ni;
    p->h_sd[s].coef[2] = A + (int64_t)nd * (ni + nj);
  }

  // --- PR and R tables ------------------------------------------------------
  std::vector<int32_t> pr_sdoms;
  p->h_pr.resize(p->npr);
  p->pr_gset.resize(p->npr);
  p->pr_r.resize(p->npr);
  for (int c = 0; c < p->npr; ++c) {
    DevPR &H = p->h_pr[c];
    memset(&H, 0, sizeof(H));
    const kr_sdom_desc &S = p->sdoms[pr_rep[c]];
    H.phi = p->prs[c].phi;
    H.phi_out = p->prs[c].phi_out;
    H.sig = p->prs[c].sigt_zonal;
    H.ng = ng;
    H.nz = (int32_t)nz;
    H.r = S.r_chunk;
    H.gset = S.gset;
    H.glo = S.group_lower;
    H.sdom_begin = (int32_t)pr_sdoms.size();
    for (int s = 0; s < p->nsdom; ++s)
      if (p->sdoms[s].pr_chunk == c) pr_sdoms.push_back(s);
    H.sdom_count = (int32_t)pr_sdoms.size() - H.sdom_begin;
    Str3 m = flux_strides(p->nesting, p->M, ng, nz);
    H.ma = m.a; H.mg = m.g; H.mz = m.z;
    p->pr_gset[c] = S.gset;
    p->pr_r[c] = S.r_chunk;
  }
  std::vector<DevR> hr(p->nr);
  p->max_gz = (int64_t)ng * nz;
  p->max_g = ng;
  p->max_z = nz;
  p->max_gmix = 0;
  for (int c = 0; c < p->nr; ++c) {
    const kr_r_desc &R = p->rs[c];
    DevR &H = hr[c];
    H.zone_to_mixelem = R.zone_to_mixelem;
    H.zone_to_num = R.zone_to_num_mixelem;
    H.mix_zone = R.mixelem_to_zone;
    H.mix_mat = R.mixelem_to_material;
    H.mix_frac = R.mixelem_to_fraction;
    H.volume = R.volume;
    H.vol_info = p->d_vol_info + 2 * c;
    H.num_mix = R.num_mixelem;
    H.nz = (int32_t)nz;
    p->max_gmix = std::max<int64_t>(p->max_gmix, (int64_t)ng * R.num_mixelem);
  }

  // --- sigs structure -------------------------------------------------------
  {
    const int64_t L1 = p->L1, G = p->G, nmat = p->nmat;
    const bool n_first = (p->nesting == KR_DGZ || p->nesting == KR_DZG || p->nesting == KR_ZDG);
    if (n_first) { p->sigs_str[0] = L1 * G * G; p->sigs_str[1] = G * G; p->sigs_str[2] = G; p->sigs_str[3] = 1; }
    else { p->sigs_str[0] = G * G * L1; p->sigs_str[1] = 1; p->sigs_str[2] = G * L1; p->sigs_str[3] = L1; }
    std::vector<int32_t> flags(SIGS_FLAG_BLOCKS, 0);
    if ((rc = upload(ctx, &p->d_sigs_flags, flags))) { kr_problem_destroy(p); return rc; }
    {
      const int64_t nt = sigs_scratch_doubles(nmat, L1, G, p->M, p->nr, p->max_z);
      const hipError_t e = hipMalloc((void **)&p->d_sigs_t, (size_t)nt * sizeof(double));
      if (e != hipSuccess) { kr_problem_destroy(p); return fail(KR_ERR_NOMEM, std::string("hipMalloc sigs_t: ") + hipGetErrorString(e)); }
    }
  }

  // --- sweep planes tables --------------------------------------------------
  std::vector<SweepPlanes> pref(p->nsdom), ppipe(p->nsdom), padj(p->nsdom), pseed(p->nsdom), pbj(p->nsdom);
  for (int s = 0; s < p->nsdom; ++s) {
    const DevSdom &H = p->h_sd[s];
    for (int dim = 0; dim < 3; ++dim) {
      pref[s].in[dim] = H.plane[dim];
      pref[s].out[dim] = H.plane[dim];
      ppipe[s].in[dim] = H.up[dim] >= 0 ? H.plane[dim] : nullptr;
      ppipe[s].out[dim] = H.down[dim] >= 0 ? p->h_sd[H.down[dim]].plane[dim] : H.plane[dim];
      // the adjoint sweep runs against the flow: its faces enter from the
      // forward-downwind neighbour and leave toward the forward-upwind one
      padj[s].in[dim] = H.down[dim] >= 0 ? H.plane[dim] : nullptr;
      padj[s].out[dim] = H.up[dim] >= 0 ? p->h_sd[H.up[dim]].plane[dim] : H.plane[dim];
      // seeded: where the forward flow leaves the domain, the adjoint inflow
      // is read from the own plane chunk (the caller's exiting-face cotangent)
      pseed[s].in[dim] = H.plane[dim];
      pseed[s].out[dim] = padj[s].out[dim];
      pbj[s].in[dim] = nullptr;
      pbj[s].out[dim] = H.plane[dim];
    }
  }

  // --- sweep schedules --------------------------------------------------------
  bool aligned16 = true;
  for (const DevSdom &H : p->h_sd) {
    const void *ptrs[] = {H.psi, H.rhs, H.sig, H.plane[0], H.plane[1], H.plane[2]};
    for (const void *q : ptrs) aligned16 = aligned16 && ((uintptr_t)q % 16 == 0);
  }
  p->aligned16 = aligned16;
  {
    SweepTune t;
    t.lds_max = ctx->lds_max;
    p->shape_ref = choose_sweep_shape(p->nesting, nd, ng, ni, nj, nk, 1, aligned16, t);
    p->shape_pipe = choose_sweep_shape(p->nesting, nd, ng, ni, nj, nk, 0, aligned16, t);
  }
  const int32_t pgs = (int32_t)(((int64_t)nd * ng) / p->shape_pipe.Q);
  std::vector<SweepUnit> uref, upipe;
  {
    const int32_t R = pick_R(pgs, p->shape_ref);
    for (int s = 0; s < p->nsdom; ++s) {
      p->ref_unit_begin.push_back((int32_t)uref.size());
      append_units(uref, s, pgs, R);
      p->ref_unit_count.push_back((int32_t)uref.size() - p->ref_unit_begin.back());
      order_units_xcd(p, p->shape_ref, uref, p->ref_unit_begin.back(), p->ref_unit_count.back());
    }
    std::vector<int32_t> all(p->nsdom);
    for (int s = 0; s < p->nsdom; ++s) all[s] = s;
    std::vector<std::vector<int32_t>> levels;
    if ((rc = sweep_levels(p, all, levels))) { kr_problem_destroy(p); return rc; }
    for (auto &lv : levels) {
      const int32_t Rl = pick_R((int64_t)pgs * lv.size(), p->shape_pipe);
      p->pipe_level_begin.push_back((int32_t)upipe.size());
      for (int32_t s : lv) append_units(upipe, s, pgs, Rl);
      p->pipe_level_count.push_back((int32_t)upipe.size() - p->pipe_level_begin.back());
      order_units_xcd(p, p->shape_pipe, upipe, p->pipe_level_begin.back(), p->pipe_level_count.back());
    }
  }

  if (p->shape_ref.kind == 2 || p->shape_pipe.kind == 2)
    return bad("sweep: directions per subdomain x (ni+nj+nk) exceed the LDS coefficient table");
  if ((rc = check_unit_spans(p, uref, p->shape_ref)) || (rc = check_unit_spans(p, upipe, p->shape_pipe))) {
    kr_problem_destroy(p);
    return rc;
  }
  std::vector<DevSdom> hsd = p->h_sd;
  std::vector<char> popb;
  p->pop_nblocks = population_block_table(p->h_sd.data(), p->nsdom, popb);
  if ((rc = upload(ctx, &p->d_sd, hsd)) || (rc = upload(ctx, &p->d_pr, p->h_pr)) || (rc = upload(ctx, &p->d_r, hr)) ||
      (rc = upload(ctx, &p->d_pr_sdoms, pr_sdoms)) || (rc = upload(ctx, &p->d_planes_ref, pref)) ||
      (rc = upload(ctx, &p->d_planes_pipe, ppipe)) || (rc = upload(ctx, &p->d_planes_adj, padj)) ||
      (rc = upload(ctx, &p->d_planes_adj_seed, pseed)) ||
      (rc = upload(ctx, &p->d_planes_bj, pbj)) ||
      (rc = upload(ctx, &p->d_units_ref, uref)) ||
      (rc = upload(ctx, &p->d_units_pipe, upipe))) {
    kr_problem_destroy(p);
    return rc;
  }
  // the tables and their flags as the fields hold them now (every sweep
  // entry point refreshes them before it sweeps; the autotuner's timing
  // sweeps read these)
  if ((rc = launch_sweep_coef(ctx->stream, p->d_sd, 0, p->nsdom))) {
    kr_problem_destroy(p);
    return rc;
  }
  {
    char *db = nullptr;
    if ((rc = upload(ctx, &db, popb))) { kr_problem_destroy(p); return rc; }
    p->d_pop_blocks = db;
    if (hipMalloc((void **)&p->d_pop_partial, sizeof(double) * std::max<int64_t>(1, p->pop_nblocks)) != hipSuccess ||
        hipMalloc((void **)&p->d_pop_out, sizeof(double)) != hipSuccess) {
      kr_problem_destroy(p);
      return fail(KR_ERR_NOMEM, "kr_problem_create: population buffers");
    }
  }
  *out = p;
  return KR_OK;
}

int kr_problem_destroy(kr_problem *p) {
  if (!p) return KR_OK;
  if (p->ctx) hipStreamSynchronize(p->ctx->stream);
  for (kr_batch *b : p->batches) batch_forget_problem(b, p);
  p->batches.clear();
  for (auto &tc : p->pending) {
    hipEventDestroy(tc.start);
    hipEventDestroy(tc.stop);
  }
  void *bufs[] = {p->d_sd,          p->d_pr,        p->d_r,          p->d_pr_sdoms,  p->d_coef,
                  p->d_planes_ref,  p->d_planes_pipe, p->d_units_ref, p->d_units_pipe, p->d_units_tmp,
                  p->d_pop_blocks,  p->d_pop_partial, p->d_pop_out,  p->d_sigs_flags, p->d_pop_units,
                  p->d_planes_adj,  (void *)p->d_ptrs, p->d_sigs_t, p->d_planes_bj, p->d_planes_adj_seed,
                  p->d_vol_info};
  for (void *b : bufs)
    if (b) hipFree(b);
  delete p;
  return KR_OK;
}

int kr_kconst(kr_context *ctx, double *chunk, int64_t n, double value) {
  if (!ctx || (!chunk && n > 0) || n < 0) return fail(KR_ERR_INVALID, "kr_kconst: bad argument");
  return launch_fill(ctx->stream, chunk, n, value);
}

int kr_kcopy(kr_context *ctx, double *dst, const double *src, int64_t n) {
  if (!ctx || ((!dst || !src) && n > 0) || n < 0) return fail(KR_ERR_INVALID, "kr_kcopy: bad argument");
  return launch_copy(ctx->stream, dst, src, n);
}

int kr_ltimes(kr_problem *p, int32_t accumulate) {
  if (!p) return fail(KR_ERR_INVALID, "kr_ltimes: null problem");
  CallTimer t(p, "LTimes");
  return launch_ltimes(launch_ctx(p), accumulate ? 1 : 0);
}

int kr_lplustimes(kr_problem *p, int32_t accumulate) {
  if (!p) return fail(KR_ERR_INVALID, "kr_lplustimes: null problem");
  CallTimer t(p, "LPlusTimes");
  return launch_lplustimes(launch_ctx(p), accumulate ? 1 : 0);
}

int kr_scattering(kr_problem *p, int32_t accumulate) {
  if (!p) return fail(KR_ERR_INVALID, "kr_scattering: null problem");
  CallTimer t(p, "Scattering");
  return launch_scattering(launch_ctx(p), p->pr_gset.data(), p->pr_r.data(), accumulate ? 1 : 0);
}

int kr_source(kr_problem *p, double strength) {
  if (!p) return fail(KR_ERR_INVALID, "kr_source: null problem");
  CallTimer t(p, "Source");
  return launch_source(launch_ctx(p), strength);
}

int kr_sweep_subdomain(kr_problem *p, int32_t sdom) {
  if (!p) return fail(KR_ERR_INVALID, "kr_sweep_subdomain: null problem");
  if (sdom < 0 || sdom >= p->nsdom) return fail(KR_ERR_INVALID, "kr_sweep_subdomain: bad subdomain id");
  CallTimer t(p, "SweepSubdomain");
  int rc = launch_sweep_coef(p->ctx->stream, p->d_sd, sdom, 1);
  if (rc) return rc;
  return launch_sweep(p->ctx->stream, p->d_sd, p->d_planes_ref, p->d_units_ref + p->ref_unit_begin[sdom],
                      p->ref_unit_count[sdom], p->shape_ref);
}

// SweepSolver body; pop (zone-fastest nestings only) receives one population
// partial per sweep unit, indexed like the units of the launches.
static int sweep_solver_impl(kr_problem *p, const int32_t *list, int32_t num, double **pop_units, int64_t *npop) {
  {
    int rc = launch_sweep_coef(p->ctx->stream, p->d_sd, 0, p->nsdom);
    if (rc) return rc;
  }
  bool all = (num == p->nsdom);
  for (int32_t i = 0; all && i < num; ++i) all = (list[i] == i);
  if (all) {
    double *pop = nullptr;
    if (pop_units) {
      const int64_t n = p->pipe_level_begin.empty() ? 0
                                                    : (int64_t)p->pipe_level_begin.back() + p->pipe_level_count.back();
      if (n > p->pop_units_cap) {
        // pointer and capacity are cleared first: a failed allocation leaves
        // no freed buffer behind a non-zero capacity
        if (p->d_pop_units) KR_HIP_CHECK(hipFree(p->d_pop_units));
        p->d_pop_units = nullptr;
        p->pop_units_cap = 0;
        KR_HIP_CHECK(hipMalloc((void **)&p->d_pop_units, std::max<int64_t>(n, 1) * sizeof(double)));
        p->pop_units_cap = n;
      }
      pop = p->d_pop_units;
      *pop_units = pop;
      *npop = n;
      int rc = launch_vol_info(p->ctx->stream, p->d_r, p->nr);
      if (rc) return rc;
    }
    for (size_t lv = 0; lv < p->pipe_level_begin.size(); ++lv) {
      CallTimer tk(p, "k_sweep");  // one sweep-kernel launch (one zone-set stage)
      int rc = launch_sweep(p->ctx->stream, p->d_sd, p->d_planes_pipe, p->d_units_pipe + p->pipe_level_begin[lv],
                            p->pipe_level_count[lv], p->shape_pipe, pop ? pop + p->pipe_level_begin[lv] : nullptr);
      if (rc) return rc;
    }
    return KR_OK;
  }
  if (pop_units) return fail(KR_ERR_INVALID, "sweep: fused population needs the identity subdomain list");
  std::vector<int32_t> v(list, list + num);
  std::vector<std::vector<int32_t>> levels;
  int rc = sweep_levels(p, v, levels);
  if (rc) return rc;
  const int32_t pgs = (int32_t)(((int64_t)p->h_sd[0].nd * p->h_sd[0].ng) / p->shape_pipe.Q);
  std::vector<SweepUnit> u;
  std::vector<int32_t> lb, lc;
  for (auto &lv : levels) {
    const int32_t Rl = pick_R((int64_t)pgs * lv.size(), p->shape_pipe);
    lb.push_back((int32_t)u.size());
    for (int32_t s : lv) append_units(u, s, pgs, Rl);
    lc.push_back((int32_t)u.size() - lb.back());
    order_units_xcd(p, p->shape_pipe, u, lb.back(), lc.back());
  }
  if ((rc = check_unit_spans(p, u, p->shape_pipe))) return rc;
  KR_HIP_CHECK(hipStreamSynchronize(p->ctx->stream));
  if ((int32_t)u.size() > p->units_tmp_cap) {
    if (p->d_units_tmp) KR_HIP_CHECK(hipFree(p->d_units_tmp));
    p->d_units_tmp = nullptr;
    p->units_tmp_cap = 0;
    KR_HIP_CHECK(hipMalloc((void **)&p->d_units_tmp, u.size() * sizeof(SweepUnit)));
    p->units_tmp_cap = (int32_t)u.size();
  }
  KR_HIP_CHECK(hipMemcpy(p->d_units_tmp, u.data(), u.size() * sizeof(SweepUnit), hipMemcpyHostToDevice));
  for (size_t l = 0; l < lb.size(); ++l) {
    rc = launch_sweep(p->ctx->stream, p->d_sd, p->d_planes_pipe, p->d_units_tmp + lb[l], lc[l], p->shape_pipe);
    if (rc) return rc;
  }
  return KR_OK;
}

int kr_sweep_solver(kr_problem *p, const int32_t *list, int32_t num) {
  if (!p || (!list && num > 0) || num < 0) return fail(KR_ERR_INVALID, "kr_sweep_solver: bad argument");
  CallTimer t(p, "SweepSolver");
  return sweep_solver_impl(p, list, num, nullptr, nullptr);
}

// SweepSolver with BlockJacobiComm on one rank (SweepSolver.cpp:23-87,
// BlockJacobiComm.cpp:23-116): every plane chunk starts from zero (the
// SweepSolver fill plus the communicator's fresh old planes, copied in by the
// reversed kCopy and the on-rank sends), so each listed subdomain is swept
// with zero incoming faces, independently of the others: one launch.
int kr_sweep_solver_bj(kr_problem *p, const int32_t *list, int32_t num) {
  if (!p || (!list && num > 0) || num < 0) return fail(KR_ERR_INVALID, "kr_sweep_solver_bj: bad argument");
  std::vector<char> listed(p->nsdom, 0);
  for (int32_t i = 0; i < num; ++i) {
    if (list[i] < 0 || list[i] >= p->nsdom || listed[list[i]]++)
      return fail(KR_ERR_INVALID, "kr_sweep_solver_bj: subdomain ids must be valid and distinct");
  }
  CallTimer t(p, "SweepSolver");
  // planes of unlisted subdomains end as the zero old planes; listed ones
  // are overwritten by their own outgoing faces
  for (int s = 0; s < p->nsdom; ++s) {
    if (listed[s]) continue;
    const kr_sdom_desc &S = p->sdoms[s];
    const int64_t dg = (int64_t)S.num_directions * S.num_groups;
    const int64_t n[3] = {dg * S.nj * S.nk, dg * S.ni * S.nk, dg * S.ni * S.nj};
    for (int dim = 0; dim < 3; ++dim) {
      int rc = launch_fill(p->ctx->stream, p->h_sd[s].plane[dim], n[dim], 0.0);
      if (rc) return rc;
    }
  }
  if (num == 0) return KR_OK;
  int rc = launch_sweep_coef(p->ctx->stream, p->d_sd, 0, p->nsdom);
  if (rc) return rc;
  const int32_t pgs = (int32_t)(((int64_t)p->h_sd[0].nd * p->h_sd[0].ng) / p->shape_pipe.Q);
  const int32_t R = pick_R((int64_t)pgs * num, p->shape_pipe);
  std::vector<SweepUnit> u;
  for (int32_t i = 0; i < num; ++i) append_units(u, list[i], pgs, R);
  order_units_xcd(p, p->shape_pipe, u, 0, u.size());
  if ((rc = check_unit_spans(p, u, p->shape_pipe))) return rc;
  KR_HIP_CHECK(hipStreamSynchronize(p->ctx->stream));
  if ((int32_t)u.size() > p->units_tmp_cap) {
    if (p->d_units_tmp) KR_HIP_CHECK(hipFree(p->d_units_tmp));
    p->d_units_tmp = nullptr;
    p->units_tmp_cap = 0;
    KR_HIP_CHECK(hipMalloc((void **)&p->d_units_tmp, u.size() * sizeof(SweepUnit)));
    p->units_tmp_cap = (int32_t)u.size();
  }
  KR_HIP_CHECK(hipMemcpy(p->d_units_tmp, u.data(), u.size() * sizeof(SweepUnit), hipMemcpyHostToDevice));
  CallTimer tk(p, "k_sweep");
  return launch_sweep(p->ctx->stream, p->d_sd, p->d_planes_bj, p->d_units_tmp, (int32_t)u.size(), p->shape_pipe);
}

int kr_sweep_solver_population(kr_problem *p, const int32_t *list, int32_t num, double *population) {
  if (!p || !list || !population || num != p->nsdom)
    return fail(KR_ERR_INVALID, "kr_sweep_solver_population: the list must name every subdomain once");
  std::vector<int32_t> seen(p->nsdom, 0);
  for (int32_t i = 0; i < num; ++i) {
    if (list[i] < 0 || list[i] >= p->nsdom || seen[list[i]]++)
      return fail(KR_ERR_INVALID, "kr_sweep_solver_population: the list must name every subdomain once");
  }
  // the identity order is the pipeline's own schedule; any other complete
  // list sweeps the same dependency levels, so it is swept in that order
  std::vector<int32_t> ident(p->nsdom);
  for (int32_t s = 0; s < p->nsdom; ++s) ident[s] = s;
  {
    double *pu = nullptr;
    int64_t n = 0;
    {
      CallTimer t(p, "SweepSolver");
      int rc = sweep_solver_impl(p, ident.data(), p->nsdom, &pu, &n);
      if (rc) return rc;
    }
    {
      CallTimer t(p, "Population");
      int rc = launch_sum(p->ctx->stream, pu, n, p->d_pop_out);
      if (rc) return rc;
    }
    KR_HIP_CHECK(hipMemcpyAsync(population, p->d_pop_out, sizeof(double), hipMemcpyDeviceToHost, p->ctx->stream));
    KR_HIP_CHECK(hipStreamSynchronize(p->ctx->stream));
    return KR_OK;
  }
}

// ---- adjoint (transposed) operators and parameter gradients ---------------

// Device copy of a per-PR-chunk pointer array (re-uploaded only when it
// changes; the stream is drained first so no running kernel reads it).
static int put_ptrs(kr_problem *p, int slot, const void *const *ptrs, const double **&dev) {
  std::vector<const double *> v(p->npr);
  for (int c = 0; c < p->npr; ++c) {
    if (!ptrs[c]) return fail(KR_ERR_INVALID, "null chunk pointer");
    v[c] = (const double *)ptrs[c];
  }
  const int n = std::max(1, p->npr);
  if (!p->d_ptrs) {
    KR_HIP_CHECK(hipMalloc((void **)&p->d_ptrs, sizeof(double *) * n * kr_problem::PTR_SLOTS));
  }
  dev = p->d_ptrs + (size_t)slot * n;
  if (v == p->ptrs_last[slot]) return KR_OK;
  KR_HIP_CHECK(hipStreamSynchronize(p->ctx->stream));
  KR_HIP_CHECK(hipMemcpy(dev, v.data(), sizeof(double *) * p->npr, hipMemcpyHostToDevice));
  p->ptrs_last[slot] = v;
  return KR_OK;
}

static void sigt_layout(const kr_problem *p, int64_t *s_g, int64_t *s_z) {
  const DevSdom &S = p->h_sd[0];
  sigt_strides(p->nesting, S.ng, S.nz, s_g, s_z);
}

int kr_ltimes_adjoint(kr_problem *p) {
  if (!p) return fail(KR_ERR_INVALID, "kr_ltimes_adjoint: null problem");
  CallTimer t(p, "LTimesAdjoint");
  return launch_ltimes_adjoint(launch_ctx(p));
}

int kr_lplustimes_adjoint(kr_problem *p) {
  if (!p) return fail(KR_ERR_INVALID, "kr_lplustimes_adjoint: null problem");
  CallTimer t(p, "LPlusTimesAdjoint");
  return launch_lplustimes_adjoint(launch_ctx(p));
}

int kr_scattering_adjoint(kr_problem *p, int32_t accumulate) {
  if (!p) return fail(KR_ERR_INVALID, "kr_scattering_adjoint: null problem");
  CallTimer t(p, "ScatteringAdjoint");
  return launch_scattering(launch_ctx(p), p->pr_gset.data(), p->pr_r.data(), accumulate ? 1 : 0, true);
}

int kr_scattering_zonal(kr_problem *p, const double *const *sigs_zonal, int32_t accumulate) {
  if (!p || !sigs_zonal) return fail(KR_ERR_INVALID, "kr_scattering_zonal: null argument");
  const double **dev = nullptr;
  int rc = put_ptrs(p, 0, (const void *const *)sigs_zonal, dev);
  if (rc) return rc;
  CallTimer t(p, "Scattering");
  int64_t sg, sz;
  sigt_layout(p, &sg, &sz);
  return launch_scattering_zonal(launch_ctx(p), dev, sg, sz, accumulate ? 1 : 0, false);
}

int kr_scattering_zonal_adjoint(kr_problem *p, const double *const *sigs_zonal, int32_t accumulate) {
  if (!p || !sigs_zonal) return fail(KR_ERR_INVALID, "kr_scattering_zonal_adjoint: null argument");
  const double **dev = nullptr;
  int rc = put_ptrs(p, 0, (const void *const *)sigs_zonal, dev);
  if (rc) return rc;
  CallTimer t(p, "ScatteringAdjoint");
  int64_t sg, sz;
  sigt_layout(p, &sg, &sz);
  return launch_scattering_zonal(launch_ctx(p), dev, sg, sz, accumulate ? 1 : 0, true);
}

static int sweep_solver_adjoint_on(kr_problem *p, const SweepPlanes *planes) {
  CallTimer t(p, "SweepSolverAdjoint");
  int rc = launch_sweep_coef(p->ctx->stream, p->d_sd, 0, p->nsdom);
  if (rc) return rc;
  for (size_t n = p->pipe_level_begin.size(); n-- > 0;) {  // the forward stages in reverse order
    CallTimer tk(p, "k_sweep_adjoint");
    rc = launch_sweep(p->ctx->stream, p->d_sd, planes, p->d_units_pipe + p->pipe_level_begin[n],
                      p->pipe_level_count[n], p->shape_pipe, nullptr, true);
    if (rc) return rc;
  }
  return KR_OK;
}

int kr_sweep_solver_adjoint(kr_problem *p) {
  if (!p) return fail(KR_ERR_INVALID, "kr_sweep_solver_adjoint: null problem");
  return sweep_solver_adjoint_on(p, p->d_planes_adj);
}

int kr_sweep_solver_adjoint_seeded(kr_problem *p) {
  if (!p) return fail(KR_ERR_INVALID, "kr_sweep_solver_adjoint_seeded: null problem");
  return sweep_solver_adjoint_on(p, p->d_planes_adj_seed);
}

int kr_sdom_exits(kr_problem *p, int32_t sdom, int32_t *exits) {
  if (!p || !exits || sdom < 0 || sdom >= p->nsdom) return fail(KR_ERR_INVALID, "kr_sdom_exits: bad argument");
  for (int dim = 0; dim < 3; ++dim) exits[dim] = p->h_sd[sdom].down[dim] < 0 ? 1 : 0;
  return KR_OK;
}

static int same_shape(const kr_problem *a, const kr_problem *b) {
  if (!a || !b) return fail(KR_ERR_INVALID, "null problem");
  if (a->ctx != b->ctx) return fail(KR_ERR_INVALID, "forward and adjoint problems must share a context");
  if (a->nesting != b->nesting || a->nsdom != b->nsdom || a->npr != b->npr || a->M != b->M)
    return fail(KR_ERR_INVALID, "forward and adjoint problems differ in shape");
  for (int s = 0; s < a->nsdom; ++s) {
    const DevSdom &x = a->h_sd[s], &y = b->h_sd[s];
    if (x.nd != y.nd || x.ng != y.ng || x.ni != y.ni || x.nj != y.nj || x.nk != y.nk || x.pr != y.pr)
      return fail(KR_ERR_INVALID, "forward and adjoint problems differ in shape");
  }
  return KR_OK;
}

int kr_grad_sigt(kr_problem *fwd, kr_problem *adj, double *const *grad, int32_t accumulate) {
  int rc = same_shape(fwd, adj);
  if (rc) return rc;
  if (!grad) return fail(KR_ERR_INVALID, "kr_grad_sigt: null gradient array");
  const double **dev = nullptr;
  if ((rc = put_ptrs(fwd, 1, (const void *const *)grad, dev))) return rc;
  CallTimer t(fwd, "GradSigt");
  int64_t sg, sz;
  sigt_layout(fwd, &sg, &sz);
  return launch_grad_sigt(launch_ctx(fwd), adj->d_sd, (double *const *)dev, sg, sz, accumulate ? 1 : 0);
}

int kr_grad_sigs_zonal(kr_problem *fwd, kr_problem *adj, double *const *grad, int32_t accumulate) {
  int rc = same_shape(fwd, adj);
  if (rc) return rc;
  if (!grad) return fail(KR_ERR_INVALID, "kr_grad_sigs_zonal: null gradient array");
  const double **dev = nullptr;
  if ((rc = put_ptrs(fwd, 2, (const void *const *)grad, dev))) return rc;
  CallTimer t(fwd, "GradSigs");
  int64_t sg, sz;
  sigt_layout(fwd, &sg, &sz);
  return launch_grad_sigs_zonal(launch_ctx(fwd), adj->d_pr, (double *const *)dev, sg, sz, accumulate ? 1 : 0);
}

// ---- autotuner ---------------------------------------------------------------

// Pipelined-sweep schedule for a shape and unit-size scale: units grouped
// by dependency level.
static int build_pipe_schedule(kr_problem *p, const SweepShape &sh, int r_scale_pct, std::vector<SweepUnit> &units,
                               std::vector<int32_t> &lbeg, std::vector<int32_t> &lcnt) {
  std::vector<int32_t> all(p->nsdom);
  for (int s = 0; s < p->nsdom; ++s) all[s] = s;
  std::vector<std::vector<int32_t>> levels;
  int rc = sweep_levels(p, all, levels);
  if (rc) return rc;
  const int32_t pgs = (int32_t)(((int64_t)p->h_sd[0].nd * p->h_sd[0].ng) / sh.Q);
  units.clear();
  lbeg.clear();
  lcnt.clear();
  for (auto &lv : levels) {
    int64_t R = (int64_t)pick_R((int64_t)pgs * lv.size(), sh) * r_scale_pct / 100;
    R = std::max<int64_t>(1, std::min<int64_t>(R, 64));
    lbeg.push_back((int32_t)units.size());
    for (int32_t s : lv) append_units(units, s, pgs, (int32_t)R);
    lcnt.push_back((int32_t)units.size() - lbeg.back());
    order_units_xcd(p, sh, units, lbeg.back(), lcnt.back());
  }
  return check_unit_spans(p, units, sh);
}

static std::vector<int64_t> tune_key(const kr_problem *p) {
  const DevSdom &S = p->h_sd[0];
  return {p->nesting, p->nsdom, S.nd, S.ng, S.ni, S.nj, S.nk, p->aligned16 ? 1 : 0};
}

static int install_choice(kr_problem *p, const TuneChoice &c) {
  const DevSdom &S = p->h_sd[0];
  SweepTune t;
  t.v_max = c.v_max;
  t.cap_div = c.cap_div;
  t.lds_max = p->ctx->lds_max;
  const SweepShape sh = choose_sweep_shape(p->nesting, S.nd, S.ng, S.ni, S.nj, S.nk, 0, p->aligned16, t);
  if (sh.kind == 2) return fail(KR_ERR_UNSUPPORTED, "autotune: shape does not fit");
  std::vector<SweepUnit> u;
  std::vector<int32_t> lb, lc;
  int rc = build_pipe_schedule(p, sh, c.r_scale_pct, u, lb, lc);
  if (rc) return rc;
  KR_HIP_CHECK(hipStreamSynchronize(p->ctx->stream));
  SweepUnit *d = nullptr;
  if ((rc = upload(p->ctx, &d, u))) return rc;
  if (p->d_units_pipe) KR_HIP_CHECK(hipFree(p->d_units_pipe));
  p->d_units_pipe = d;
  p->pipe_level_begin = lb;
  p->pipe_level_count = lc;
  p->shape_pipe = sh;
  p->tune = c;
  ++p->sched_version;
  return KR_OK;
}

int kr_problem_set_tuning(kr_problem *p, int32_t v_max, int32_t cap_div, int32_t r_scale_pct) {
  if (!p || v_max < 0 || v_max > 4 || cap_div < 1 || cap_div > 16 || r_scale_pct < 1 || r_scale_pct > 10000)
    return fail(KR_ERR_INVALID, "kr_problem_set_tuning: bad argument");
  TuneChoice c;
  c.v_max = v_max;
  c.cap_div = cap_div;
  c.r_scale_pct = r_scale_pct;
  return install_choice(p, c);
}

int kr_problem_sweep_kind(kr_problem *p, int32_t *kind) {
  if (!p || !kind) return fail(KR_ERR_INVALID, "kr_problem_sweep_kind: null argument");
  *kind = (p->shape_pipe.kind == 0 && p->shape_pipe.kring) ? 4 : p->shape_pipe.kind;
  return KR_OK;
}

int kr_problem_autotune(kr_problem *p, int32_t *choice) {
  if (!p) return fail(KR_ERR_INVALID, "kr_problem_autotune: null problem");
  const std::vector<int64_t> key = tune_key(p);
  auto it = p->ctx->tuned.find(key);
  TuneChoice best;
  if (it != p->ctx->tuned.end()) {
    best = it->second;
  } else {
    std::vector<TuneChoice> cand;
    const bool zfast = p->nesting == KR_DGZ || p->nesting == KR_GDZ;
    for (int pct : {50, 100, 200}) {
      if (zfast) {
        // two zones per lane (the default), one, or four
        for (int v : {0, 1, 4}) {
          TuneChoice c;
          c.v_max = v;
          c.r_scale_pct = pct;
          cand.push_back(c);
        }
      } else {
        // the systolic kernel with its largest patch or half of it (v_max 0),
        // the wave kernel with 16 or 8 waves per workgroup (v_max 2), the
        // systolic kernel with its patches in k-fastest order and the faces
        // between k-patches in an LDS ring, 1024 or 512 threads (v_max 3);
        // shapes a kernel does not cover fall back to the plain systolic one
        for (int v : {0, 2, 3})
          for (int cd : {1, 2}) {
            TuneChoice c;
            c.v_max = v;
            c.cap_div = cd;
            c.r_scale_pct = pct;
            cand.push_back(c);
          }
      }
    }
    hipEvent_t e0, e1;
    KR_HIP_CHECK(hipEventCreate(&e0));
    KR_HIP_CHECK(hipEventCreate(&e1));
    const bool verbose = getenv("KR_AUTOTUNE_VERBOSE") != nullptr;  // candidate timings on stderr
    double best_ms = 1e300;
    int rc = KR_OK;
    for (TuneChoice &c : cand) {
      if ((rc = install_choice(p, c))) {
        if (rc == KR_ERR_UNSUPPORTED) continue;
        break;
      }
      // timed as the pipelined solve runs it: with the fused population
      // partials (one per unit) and the volume check before it
      const int64_t nu = p->pipe_level_begin.empty() ? 0
                                                     : (int64_t)p->pipe_level_begin.back() + p->pipe_level_count.back();
      if (nu > p->pop_units_cap) {
        if (p->d_pop_units) KR_HIP_CHECK(hipFree(p->d_pop_units));
        p->d_pop_units = nullptr;
        p->pop_units_cap = 0;
        KR_HIP_CHECK(hipMalloc((void **)&p->d_pop_units, std::max<int64_t>(nu, 1) * sizeof(double)));
        p->pop_units_cap = nu;
      }
      if ((rc = launch_vol_info(p->ctx->stream, p->d_r, p->nr))) break;
      double ms = 1e300;
      for (int rep = 0; rep < 3 && !rc; ++rep) {  // first run warms up; keep the best of the rest
        hipEventRecord(e0, p->ctx->stream);
        for (size_t lv = 0; lv < p->pipe_level_begin.size() && !rc; ++lv)
          rc = launch_sweep(p->ctx->stream, p->d_sd, p->d_planes_pipe, p->d_units_pipe + p->pipe_level_begin[lv],
                            p->pipe_level_count[lv], p->shape_pipe, p->d_pop_units + p->pipe_level_begin[lv]);
        hipEventRecord(e1, p->ctx->stream);
        hipEventSynchronize(e1);
        float t = 0.f;
        hipEventElapsedTime(&t, e0, e1);
        if (rep > 0) ms = std::min(ms, (double)t);
      }
      c.ms = ms;
      if (verbose)
        fprintf(stderr, "kr_problem_autotune: v_max %d cap_div %d R %d%%: %.3f ms\n", c.v_max, c.cap_div,
                c.r_scale_pct, ms);
      if (!rc && ms < best_ms) {
        best_ms = ms;
        best = c;
      }
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (rc) return rc;
    p->ctx->tuned[key] = best;
  }
  int rc = install_choice(p, best);
  if (rc) return rc;
  if (choice) {
    choice[0] = (p->shape_pipe.kind == 0 && p->shape_pipe.kring) ? 4 : p->shape_pipe.kind;  // as kr_problem_sweep_kind
    choice[1] = p->shape_pipe.kind == 1 ? p->shape_pipe.V : p->shape_pipe.PJ * p->shape_pipe.PK;
    choice[2] = best.r_scale_pct;
  }
  return KR_OK;
}

// ---- batched persistent launch ----------------------------------------------

// A batch runs its problems' iterations as multi-problem launches: the
// tables of every problem (subdomains, PR and R chunks, planes, sweep units)
// are concatenated, with their indices shifted, and one launch covers the
// chunks or units of all problems that take the same kernel variant.  Each
// chunk's and each unit's result depends only on its own tables, so every
// value equals the problem's own launches bit for bit; the population of a
// problem adds its units' partials in its own unit order.
constexpr int BATCH_BRANCHES = 8;
constexpr int BATCH_PARTS = 4;  // problem groups advanced on their own graph branches

struct BatchSweep {
  SweepShape shape;     // the members' shape (row-scan kernel: their common RG, V, PK)
  int64_t coef = 0;     // LDS coefficient table (doubles): the largest over the members
  int32_t ubeg = 0, ucnt = 0;
};

struct BatchLt {
  int M = 0, nd = 0, dfast = 0;
  int32_t pbeg = 0, pcnt = 0;  // range of the group's reordered PR table
  int64_t max_gz = 0;
};

// The problems of one part: combined tables and launch lists.
struct BatchPart {
  int32_t p0 = 0, np = 0;            // problems [p0, p0 + np) of the batch
  char *d_tab = nullptr;             // the tables, one device block
  DevSdom *d_sd = nullptr;
  DevPR *d_pr = nullptr, *d_pr_lt = nullptr, *d_pr_lp = nullptr;
  DevR *d_r = nullptr;
  int32_t *d_pr_sdoms = nullptr;
  SweepPlanes *d_planes = nullptr;
  SweepUnit *d_units = nullptr;
  DevScat *d_scat_chunk = nullptr;
  int64_t *d_pop_list = nullptr, *d_pop_begin = nullptr;
  double *d_pop_units = nullptr;     // one population partial per combined unit
  int32_t nsd = 0, npr = 0, nr = 0;
  int64_t max_gz = 0, max_gmix = 0;
  std::vector<BatchLt> lt, lp;
  std::vector<std::vector<BatchSweep>> levels;
  hipStream_t stream = nullptr;      // the part's graph branch while capturing
  hipEvent_t done = nullptr;
};

struct kr_batch {
  kr_context *ctx = nullptr;
  std::vector<kr_problem *> probs;
  std::vector<double> strength;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int32_t graph_niter = -1;
  std::vector<int64_t> versions;     // the problems' schedule versions at capture
  std::vector<int32_t> coupled;      // per problem: its sigs couples groups or moments (at capture)
  double *d_pop = nullptr;           // [problem][iteration]
  int64_t pop_cap = 0;
  int32_t *d_flags = nullptr;        // SIGS_FLAG_BLOCKS per problem
  DevScat *d_scat_all = nullptr;     // one entry per problem (structure check)
  std::vector<BatchPart> parts;
  hipStream_t branch[BATCH_BRANCHES] = {};
  hipEvent_t join[BATCH_BRANCHES] = {};
  hipEvent_t fork = nullptr;
  bool member_destroyed = false;     // a member problem was destroyed while the batch lived
};

}  // extern "C"

static void batch_forget_problem(kr_batch *b, kr_problem *p) {
  for (kr_problem *&q : b->probs)
    if (q == p) q = nullptr;
  b->member_destroyed = true;
}

// Enqueues launch(stream, i) for i < n on BATCH_BRANCHES streams forked from
// s0 and joined back (graph branches when captured): independent launches of
// one step may overlap, so the tail of one fills with the start of another.
template <class F>
static int batch_parallel(kr_batch *b, hipStream_t s0, int n, bool branches, F &&launch) {
  if (n <= 1 || !branches) {
    for (int i = 0; i < n; ++i) {
      int rc = launch(s0, i);
      if (rc) return rc;
    }
    return KR_OK;
  }
  KR_HIP_CHECK(hipEventRecord(b->fork, s0));
  const int used = std::min(n, BATCH_BRANCHES);
  for (int k = 0; k < used; ++k) KR_HIP_CHECK(hipStreamWaitEvent(b->branch[k], b->fork, 0));
  for (int i = 0; i < n; ++i) {
    int rc = launch(b->branch[i % BATCH_BRANCHES], i);
    if (rc) return rc;
  }
  for (int k = 0; k < used; ++k) {
    KR_HIP_CHECK(hipEventRecord(b->join[k], b->branch[k]));
    KR_HIP_CHECK(hipStreamWaitEvent(s0, b->join[k], 0));
  }
  return KR_OK;
}

extern "C" {

static void batch_free_tables(kr_batch *b) {
  for (BatchPart &P : b->parts) {
    if (P.d_tab) hipFree(P.d_tab);
    if (P.d_pop_units) hipFree(P.d_pop_units);
    if (P.stream) hipStreamDestroy(P.stream);
    if (P.done) hipEventDestroy(P.done);
  }
  b->parts.clear();
  if (b->d_scat_all) hipFree(b->d_scat_all);
  b->d_scat_all = nullptr;
}

int kr_batch_destroy(kr_batch *b) {
  if (!b) return KR_OK;
  if (b->ctx) hipStreamSynchronize(b->ctx->stream);
  for (kr_problem *q : b->probs)
    if (q) q->batches.erase(std::remove(q->batches.begin(), q->batches.end(), b), q->batches.end());
  if (b->exec) hipGraphExecDestroy(b->exec);
  if (b->graph) hipGraphDestroy(b->graph);
  if (b->d_pop) hipFree(b->d_pop);
  if (b->d_flags) hipFree(b->d_flags);
  for (int k = 0; k < BATCH_BRANCHES; ++k) {
    if (b->branch[k]) hipStreamDestroy(b->branch[k]);
    if (b->join[k]) hipEventDestroy(b->join[k]);
  }
  if (b->fork) hipEventDestroy(b->fork);
  batch_free_tables(b);
  delete b;
  return KR_OK;
}

int kr_batch_create(kr_context *ctx, kr_problem *const *problems, const double *source_strength, int32_t n,
                    kr_batch **out) {
  if (!ctx || !problems || !source_strength || n <= 0 || !out)
    return fail(KR_ERR_INVALID, "kr_batch_create: bad argument");
  *out = nullptr;
  auto *b = new kr_batch();
  b->ctx = ctx;
  for (int32_t i = 0; i < n; ++i) {
    if (!problems[i] || problems[i]->ctx != ctx) {
      kr_batch_destroy(b);
      return fail(KR_ERR_INVALID, "kr_batch_create: every problem must be registered on this context");
    }
    for (int32_t j = 0; j < i; ++j)
      if (problems[j] == problems[i]) {
        kr_batch_destroy(b);
        return fail(KR_ERR_INVALID, "kr_batch_create: a problem appears twice");
      }
    b->probs.push_back(problems[i]);
    b->strength.push_back(source_strength[i]);
  }
  if (hipMalloc((void **)&b->d_flags, sizeof(int32_t) * SIGS_FLAG_BLOCKS * n) != hipSuccess) {
    b->d_flags = nullptr;
    kr_batch_destroy(b);
    return fail(KR_ERR_NOMEM, "kr_batch_create: structure flags");
  }
  bool ok = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) == hipSuccess;
  for (int k = 0; ok && k < BATCH_BRANCHES; ++k)
    ok = hipStreamCreateWithFlags(&b->branch[k], hipStreamNonBlocking) == hipSuccess &&
         hipEventCreateWithFlags(&b->join[k], hipEventDisableTiming) == hipSuccess;
  if (!ok) {
    kr_batch_destroy(b);
    return fail(KR_ERR_HIP, "kr_batch_create: stream/event creation failed");
  }
  for (kr_problem *q : b->probs) q->batches.push_back(b);
  *out = b;
  return KR_OK;
}

}  // extern "C"

// Host image of the combined tables: each part 256-byte aligned in one block.
struct TabImage {
  std::vector<char> bytes;
  template <class T>
  size_t add(const std::vector<T> &v) {
    const size_t off = (bytes.size() + 255) & ~(size_t)255;
    bytes.resize(off + v.size() * sizeof(T));
    if (!v.empty()) memcpy(bytes.data() + off, v.data(), v.size() * sizeof(T));
    return off;
  }
};

extern "C" {

// Classifies every problem's sigs into the batch's flags and returns, per
// problem, whether it couples groups or moments.
static int batch_coupling(kr_batch *b, std::vector<int32_t> &coupled) {
  const size_t n = b->probs.size();
  hipStream_t st = b->ctx->stream;
  int rc = launch_sigs_structure_multi(st, b->d_scat_all, (int)n);
  if (rc) return rc;
  std::vector<int32_t> f;
  if ((rc = download(b->ctx, f, b->d_flags, n * SIGS_FLAG_BLOCKS))) return rc;
  coupled.assign(n, 0);
  for (size_t i = 0; i < n; ++i)
    for (int k = 0; k < SIGS_FLAG_BLOCKS; ++k) coupled[i] |= f[i * SIGS_FLAG_BLOCKS + k] != 0;
  return KR_OK;
}

// Builds one part's combined tables and launch lists from its problems'
// current device tables and pipelined sweep schedules.
static int build_part(kr_batch *b, BatchPart &B) {
  kr_context *ctx = b->ctx;
  const int np = B.np;
  kr_problem *const *probs = b->probs.data() + B.p0;
  std::vector<DevSdom> sd;
  std::vector<DevPR> pr;
  std::vector<DevR> rr;
  std::vector<int32_t> prs, pr_prob;
  std::vector<SweepPlanes> planes;
  std::vector<DevScat> scat_prob, scat_chunk;
  std::vector<int32_t> sd_off(np), pr_off(np);
  // per problem: its pipelined units (combined sdom ids, PK in `reserved`)
  std::vector<std::vector<SweepUnit>> units(np);
  int rc;
  B.max_gz = B.max_gmix = 0;
  for (int i = 0; i < np; ++i) {
    kr_problem *p = probs[i];
    sd_off[i] = (int32_t)sd.size();
    pr_off[i] = (int32_t)pr.size();
    const int32_t r_off = (int32_t)rr.size(), s_off = (int32_t)prs.size();
    std::vector<DevR> hr;
    std::vector<int32_t> hprs;
    std::vector<SweepPlanes> hpl;
    int64_t npl = 0;
    for (const DevPR &P : p->h_pr) npl = std::max<int64_t>(npl, (int64_t)P.sdom_begin + P.sdom_count);
    const int64_t nu = p->pipe_level_begin.empty() ? 0 : (int64_t)p->pipe_level_begin.back() + p->pipe_level_count.back();
    if ((rc = download(ctx, hr, p->d_r, p->nr)) || (rc = download(ctx, hprs, p->d_pr_sdoms, npl)) ||
        (rc = download(ctx, hpl, p->d_planes_pipe, p->nsdom)) || (rc = download(ctx, units[i], p->d_units_pipe, nu)))
      return rc;
    for (DevSdom S : p->h_sd) {
      S.pr += pr_off[i];
      S.r += r_off;
      sd.push_back(S);
    }
    for (DevPR P : p->h_pr) {
      P.r += r_off;
      P.sdom_begin += s_off;
      pr.push_back(P);
      pr_prob.push_back(i);
      B.max_gz = std::max<int64_t>(B.max_gz, (int64_t)P.ng * P.nz);
    }
    for (const DevR &R : hr) {
      rr.push_back(R);
    }
    for (size_t k = 0; k < p->h_pr.size(); ++k)
      B.max_gmix = std::max<int64_t>(B.max_gmix, (int64_t)p->h_pr[k].ng * hr[p->h_pr[k].r].num_mix);
    for (int32_t s : hprs) prs.push_back(s + sd_off[i]);
    planes.insert(planes.end(), hpl.begin(), hpl.end());
    DevScat Q{};
    Q.sigs = p->sigs;
    for (int k = 0; k < 4; ++k) Q.str[k] = p->sigs_str[k];
    Q.nmat = p->nmat;
    Q.L1 = p->L1;
    Q.G = p->G;
    Q.strength = b->strength[B.p0 + i];
    Q.M = p->M;
    Q.problem = B.p0 + i;
    scat_prob.push_back(Q);  // flags pointer set after the upload
    for (SweepUnit &u : units[i]) u.sdom += sd_off[i];
  }
  B.nsd = (int32_t)sd.size();
  B.nr = (int32_t)rr.size();
  B.npr = (int32_t)pr.size();
  for (int32_t k = 0; k < B.npr; ++k) scat_chunk.push_back(scat_prob[pr_prob[k]]);

  // LTimes / LPlusTimes groups: chunks of equal kernel variant, reordered
  // so each group is one contiguous range
  auto group = [&](bool lplus, std::vector<DevPR> &order, std::vector<BatchLt> &out) {
    std::map<std::pair<int64_t, int>, std::vector<int32_t>> g;  // (variant, dfast) -> chunks
    std::vector<int64_t> key(np);
    std::vector<int> dfast(np);
    for (int i = 0; i < np; ++i) {
      LaunchCtx L = launch_ctx(probs[i]);
      key[i] = ltimes_variant(L, lplus);
      dfast[i] = L.dfast;
    }
    for (int32_t k = 0; k < B.npr; ++k) g[{key[pr_prob[k]], dfast[pr_prob[k]]}].push_back(k);
    for (auto &kv : g) {
      BatchLt G;
      const kr_problem *p0 = probs[pr_prob[kv.second[0]]];
      G.M = p0->M;
      G.nd = p0->h_sd[0].nd;
      G.dfast = kv.first.second;
      G.pbeg = (int32_t)order.size();
      for (int32_t k : kv.second) {
        order.push_back(pr[k]);
        G.max_gz = std::max<int64_t>(G.max_gz, (int64_t)pr[k].ng * pr[k].nz);
      }
      G.pcnt = (int32_t)kv.second.size();
      out.push_back(G);
    }
  };
  std::vector<DevPR> pr_lt, pr_lp;
  group(false, pr_lt, B.lt);
  group(true, pr_lp, B.lp);

  // sweep launches per dependency level: the row-scan kernel takes units of
  // any extents with the same row groups (RG, V) and pair order; other
  // shapes share a launch only when equal
  std::vector<SweepUnit> cu;                // combined units, launch by launch
  std::vector<std::vector<int64_t>> pos(np);  // combined position of each problem's unit
  for (int i = 0; i < np; ++i) pos[i].assign(units[i].size(), -1);
  size_t nlev = 0;
  for (int i = 0; i < np; ++i) nlev = std::max(nlev, probs[i]->pipe_level_begin.size());
  for (size_t lv = 0; lv < nlev; ++lv) {
    std::vector<std::vector<int>> groups;
    std::vector<SweepShape> gshape;
    for (int i = 0; i < np; ++i) {
      const kr_problem *p = probs[i];
      if (lv >= p->pipe_level_begin.size()) continue;
      const SweepShape &s = p->shape_pipe;
      size_t gi = 0;
      for (; gi < groups.size(); ++gi) {
        const SweepShape &t = gshape[gi];
        const bool same = s.kind == 1 ? (t.kind == 1 && t.RG == s.RG && t.V == s.V && t.PK == s.PK &&
                                         t.inner_is_g == s.inner_is_g)
                                      : memcmp(&s, &t, sizeof(SweepShape)) == 0;
        if (same) break;
      }
      if (gi == groups.size()) {
        groups.emplace_back();
        gshape.push_back(s);
      }
      groups[gi].push_back(i);
    }
    std::vector<BatchSweep> launches;
    auto emit = [&](const std::vector<int> &members, SweepShape sh) {
      BatchSweep bs;
      int64_t coef = 0;
      for (int i : members) {
        const DevSdom &S = probs[i]->h_sd[0];
        coef = std::max<int64_t>(coef, (int64_t)S.nd * (S.ni + S.nj + S.nk + 1));
      }
      const int nt = sh.kind == 1 ? sh.RG * sh.PK : sh.Q * sh.PJ * sh.PK;
      const size_t lds = sweep_lds_bytes(sh, coef);
      if (lds > ctx->lds_max || nt > 1024) return false;
      bs.shape = sh;
      bs.coef = coef;
      bs.ubeg = (int32_t)cu.size();
      for (int i : members) {
        const kr_problem *p = probs[i];
        const int32_t ub = p->pipe_level_begin[lv], uc = p->pipe_level_count[lv];
        for (int32_t u = ub; u < ub + uc; ++u) {
          pos[i][u] = (int64_t)cu.size();
          cu.push_back(units[i][u]);
        }
      }
      bs.ucnt = (int32_t)cu.size() - bs.ubeg;
      launches.push_back(bs);
      return true;
    };
    for (size_t gi = 0; gi < groups.size(); ++gi) {
      if (emit(groups[gi], gshape[gi])) continue;
      for (int i : groups[gi])  // too large together: each problem on its own
        if (!emit({i}, probs[i]->shape_pipe)) return fail(KR_ERR_UNSUPPORTED, "batch: sweep shape does not fit");
    }
    B.levels.push_back(launches);
  }
  std::vector<int64_t> plist, pbegin{0};
  for (int i = 0; i < np; ++i) {
    for (int64_t q : pos[i]) {
      if (q < 0) return fail(KR_ERR_INVALID, "batch: a sweep unit outside every level");
      plist.push_back(q);
    }
    pbegin.push_back((int64_t)plist.size());
  }

  TabImage im;
  const size_t o_sd = im.add(sd), o_pr = im.add(pr), o_lt = im.add(pr_lt), o_lp = im.add(pr_lp), o_r = im.add(rr),
               o_prs = im.add(prs), o_pl = im.add(planes), o_u = im.add(cu),
               o_sc = im.add(scat_chunk), o_l = im.add(plist), o_b = im.add(pbegin);
  KR_HIP_CHECK(hipMalloc((void **)&B.d_tab, std::max<size_t>(im.bytes.size(), 256)));
  KR_HIP_CHECK(hipMalloc((void **)&B.d_pop_units, sizeof(double) * std::max<size_t>(cu.size(), 1)));
  char *base = B.d_tab;
  B.d_sd = (DevSdom *)(base + o_sd);
  B.d_pr = (DevPR *)(base + o_pr);
  B.d_pr_lt = (DevPR *)(base + o_lt);
  B.d_pr_lp = (DevPR *)(base + o_lp);
  B.d_r = (DevR *)(base + o_r);
  B.d_pr_sdoms = (int32_t *)(base + o_prs);
  B.d_planes = (SweepPlanes *)(base + o_pl);
  B.d_units = (SweepUnit *)(base + o_u);
  B.d_scat_chunk = (DevScat *)(base + o_sc);
  B.d_pop_list = (int64_t *)(base + o_l);
  B.d_pop_begin = (int64_t *)(base + o_b);
  // flags: each chunk's problem's share of the batch's flag array
  for (int32_t k = 0; k < B.npr; ++k) {
    DevScat *q = (DevScat *)(im.bytes.data() + o_sc) + k;
    q->flags = b->d_flags + (int64_t)(B.p0 + pr_prob[k]) * SIGS_FLAG_BLOCKS;
  }
  KR_HIP_CHECK(hipMemcpyAsync(B.d_tab, im.bytes.data(), im.bytes.size(), hipMemcpyHostToDevice, ctx->stream));
  KR_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  return KR_OK;
}

// Splits the problems into consecutive parts of about equal work (unknowns),
// each advanced on its own graph branch, so that the stages of different
// parts overlap (one part's latency-bound sweeps beside another's streaming
// LTimes).  Builds every part and the per-problem structure-check table.
static int batch_build(kr_batch *b) {
  batch_free_tables(b);
  const int np = (int)b->probs.size();
  std::vector<double> work(np, 0.0);
  double total = 0.0;
  for (int i = 0; i < np; ++i) {
    for (const DevSdom &S : b->probs[i]->h_sd) work[i] += (double)S.nd * S.ng * S.nz;
    total += work[i];
  }
  const int nparts = std::max(1, std::min(BATCH_PARTS, np));
  double acc = 0.0;
  int start = 0;
  for (int k = 0; k < nparts && start < np; ++k) {
    int end = start + 1;  // at least one problem, and leave one for each later part
    acc += work[start];
    const double goal = total * (k + 1) / nparts;
    while (end < np - (nparts - k - 1) && acc + 0.5 * work[end] <= goal) acc += work[end++];
    if (k == nparts - 1)
      while (end < np) acc += work[end++];
    BatchPart part;
    part.p0 = start;
    part.np = end - start;
    b->parts.push_back(part);
    start = end;
  }
  for (BatchPart &part : b->parts) {
    int rc = build_part(b, part);
    if (rc) return rc;
    if (b->parts.size() > 1) {
      KR_HIP_CHECK(hipStreamCreateWithFlags(&part.stream, hipStreamNonBlocking));
      KR_HIP_CHECK(hipEventCreateWithFlags(&part.done, hipEventDisableTiming));
    }
  }
  std::vector<DevScat> all(np);
  for (int i = 0; i < np; ++i) {
    const kr_problem *p = b->probs[i];
    DevScat &Q = all[i];
    Q.sigs = p->sigs;
    for (int k = 0; k < 4; ++k) Q.str[k] = p->sigs_str[k];
    Q.nmat = p->nmat;
    Q.L1 = p->L1;
    Q.G = p->G;
    Q.flags = b->d_flags + (int64_t)i * SIGS_FLAG_BLOCKS;
    Q.strength = b->strength[i];
    Q.M = p->M;
    Q.problem = i;
  }
  return upload(b->ctx, &b->d_scat_all, all);
}

// One source iteration of every problem of a part (the mode-1 sequence of
// the host mirror's SteadyStateIteration), population of its problem i into
// out[i*stride].  branches: the launches of one stage on graph branches.
// Only kernel launches: capturable.
static int batch_iteration(kr_batch *b, const BatchPart &B, hipStream_t st, double *out, int64_t stride,
                           bool branches) {
  int rc;
  auto ctx_of = [&](const BatchLt &G, DevPR *tab) {
    LaunchCtx L{};
    L.stream = st;
    L.sd = B.d_sd;
    L.pr = tab + G.pbeg;
    L.r = B.d_r;
    L.pr_sdoms = B.d_pr_sdoms;
    L.num_pr = G.pcnt;
    L.num_moments = G.M;
    L.nd = G.nd;
    L.dfast = G.dfast;
    L.lds_max = b->ctx->lds_max;
    L.max_gz = G.max_gz;
    return L;
  };
  if ((rc = batch_parallel(b, st, (int)B.lt.size(), branches, [&](hipStream_t s, int i) {
         LaunchCtx L = ctx_of(B.lt[i], B.d_pr_lt);
         L.stream = s;
         return launch_ltimes(L, 0);
       })))
    return rc;
  if ((rc = launch_scatter_diag_multi(st, B.d_pr, B.d_r, B.d_scat_chunk, B.npr, B.max_gz))) return rc;
  for (int i = B.p0; i < B.p0 + B.np; ++i) {
    if (!b->coupled[i]) continue;
    LaunchCtx L = launch_ctx(b->probs[i]);
    L.stream = st;
    if ((rc = launch_scattering(L, b->probs[i]->pr_gset.data(), b->probs[i]->pr_r.data(), 0))) return rc;
  }
  if ((rc = launch_source_multi(st, B.d_pr, B.d_r, B.d_scat_chunk, B.npr, B.max_gmix))) return rc;
  if ((rc = batch_parallel(b, st, (int)B.lp.size(), branches, [&](hipStream_t s, int i) {
         LaunchCtx L = ctx_of(B.lp[i], B.d_pr_lp);
         L.stream = s;
         return launch_lplustimes(L, 0);
       })))
    return rc;
  if ((rc = launch_sweep_coef(st, B.d_sd, 0, B.nsd)) || (rc = launch_vol_info(st, B.d_r, B.nr))) return rc;
  for (const auto &lv : B.levels)
    if ((rc = batch_parallel(b, st, (int)lv.size(), branches, [&](hipStream_t s, int i) {
           const BatchSweep &W = lv[i];
           if (W.shape.kind == 1)
             return launch_sweep_rows(s, B.d_sd, B.d_planes, B.d_units + W.ubeg, W.ucnt, W.shape, W.coef,
                                      B.d_pop_units + W.ubeg, false);
           return launch_sweep(s, B.d_sd, B.d_planes, B.d_units + W.ubeg, W.ucnt, W.shape, B.d_pop_units + W.ubeg);
         })))
      return rc;
  return launch_pop_sum_multi(st, B.d_pop_units, B.d_pop_list, B.d_pop_begin, B.np, out, stride);
}

// Builds the tables and the graph of niter iterations.
static int batch_capture(kr_batch *b, int32_t niter, const std::vector<int32_t> &coupled) {
  if (b->exec) {
    KR_HIP_CHECK(hipGraphExecDestroy(b->exec));
    b->exec = nullptr;
  }
  if (b->graph) {
    KR_HIP_CHECK(hipGraphDestroy(b->graph));
    b->graph = nullptr;
  }
  const int64_t need = (int64_t)b->probs.size() * niter;
  if (need > b->pop_cap) {
    if (b->d_pop) KR_HIP_CHECK(hipFree(b->d_pop));
    b->d_pop = nullptr;
    b->pop_cap = 0;
    KR_HIP_CHECK(hipMalloc((void **)&b->d_pop, need * sizeof(double)));
    b->pop_cap = need;
  }
  b->coupled = coupled;
  hipStream_t s0 = b->ctx->stream;
  KR_HIP_CHECK(hipStreamSynchronize(s0));
  KR_HIP_CHECK(hipStreamBeginCapture(s0, hipStreamCaptureModeThreadLocal));
  int rc = KR_OK;
  if (b->parts.size() == 1) {
    for (int32_t it = 0; !rc && it < niter; ++it) rc = batch_iteration(b, b->parts[0], s0, b->d_pop + it, niter, true);
  } else {
    // each part on its own branch, its stages on that branch in order
    if (hipEventRecord(b->fork, s0) != hipSuccess) rc = fail(KR_ERR_HIP, "batch capture: fork");
    for (const BatchPart &part : b->parts) {
      if (rc) break;
      if (hipStreamWaitEvent(part.stream, b->fork, 0) != hipSuccess) {
        rc = fail(KR_ERR_HIP, "batch capture: branch");
        break;
      }
      for (int32_t it = 0; !rc && it < niter; ++it)
        rc = batch_iteration(b, part, part.stream, b->d_pop + (int64_t)part.p0 * niter + it, niter, false);
      if (!rc && (hipEventRecord(part.done, part.stream) != hipSuccess ||
                  hipStreamWaitEvent(s0, part.done, 0) != hipSuccess))
        rc = fail(KR_ERR_HIP, "batch capture: join");
    }
  }
  hipGraph_t g = nullptr;
  const hipError_t ec = hipStreamEndCapture(s0, &g);
  if (rc) {
    if (g) hipGraphDestroy(g);
    return rc;
  }
  if (ec != hipSuccess) return fail(KR_ERR_HIP, std::string("batch capture: ") + hipGetErrorString(ec));
  b->graph = g;
  KR_HIP_CHECK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  b->graph_niter = niter;
  return KR_OK;
}

int kr_batch_iterate(kr_batch *b, int32_t niter, double *populations) {
  if (!b || niter <= 0 || !populations) return fail(KR_ERR_INVALID, "kr_batch_iterate: bad argument");
  if (b->member_destroyed)
    return fail(KR_ERR_INVALID, "kr_batch_iterate: a member problem was destroyed while the batch lived");
  int rc;
  bool rebuild = b->versions.size() != b->probs.size();
  for (size_t i = 0; !rebuild && i < b->probs.size(); ++i) rebuild = b->versions[i] != b->probs[i]->sched_version;
  if (rebuild) {  // first call, or a problem re-tuned since the tables were built
    b->versions.clear();
    if ((rc = batch_build(b))) return rc;
    for (kr_problem *p : b->probs) b->versions.push_back(p->sched_version);
    b->graph_niter = -1;
  }
  // sigs as it is now decides which problems need the dense scattering
  std::vector<int32_t> coupled;
  if ((rc = batch_coupling(b, coupled))) return rc;
  if (niter != b->graph_niter || coupled != b->coupled)
    if ((rc = batch_capture(b, niter, coupled))) return rc;
  hipStream_t s0 = b->ctx->stream;
  KR_HIP_CHECK(hipGraphLaunch(b->exec, s0));
  KR_HIP_CHECK(hipMemcpyAsync(populations, b->d_pop, sizeof(double) * b->probs.size() * niter,
                              hipMemcpyDeviceToHost, s0));
  KR_HIP_CHECK(hipStreamSynchronize(s0));
  return KR_OK;
}

// kr_sweep_solver_population with the fused population's partials and sum in
// the device's static scratch array instead of problem buffers allocated on
// first use: no allocation of any kind on the device, and no timers (their
// events are created per call).  The subdomain list is checked in place.
int kr_sweep_solver_population_noalloc(kr_problem *p, const int32_t *list, int32_t num, double *population) {
  if (!p || !list || !population || num != p->nsdom)
    return fail(KR_ERR_INVALID, "kr_sweep_solver_population_noalloc: the list must name every subdomain once");
  for (int32_t i = 0; i < num; ++i) {
    if (list[i] < 0 || list[i] >= p->nsdom)
      return fail(KR_ERR_INVALID, "kr_sweep_solver_population_noalloc: the list must name every subdomain once");
    for (int32_t j = 0; j < i; ++j)
      if (list[j] == list[i])
        return fail(KR_ERR_INVALID, "kr_sweep_solver_population_noalloc: the list must name every subdomain once");
  }
  double *buf = p->ctx->noalloc_scratch;
  if (!buf) return fail(KR_ERR_HIP, "kr_sweep_solver_population_noalloc: no scratch address for this device");
  const int64_t n = p->pipe_level_begin.empty() ? 0 : (int64_t)p->pipe_level_begin.back() + p->pipe_level_count.back();
  if (n + 1 > KR_NOALLOC_SCRATCH_DOUBLES)
    return fail(KR_ERR_UNSUPPORTED, "kr_sweep_solver_population_noalloc: " + std::to_string(n) +
                                        " sweep work units need more than the " +
                                        std::to_string(KR_NOALLOC_SCRATCH_DOUBLES) + "-double scratch");
  int rc = launch_sweep_coef(p->ctx->stream, p->d_sd, 0, p->nsdom);
  if (rc) return rc;
  if ((rc = launch_vol_info(p->ctx->stream, p->d_r, p->nr))) return rc;
  // every partial is written (not accumulated) by its unit, so nothing in the
  // scratch needs clearing
  for (size_t lv = 0; lv < p->pipe_level_begin.size(); ++lv) {
    rc = launch_sweep(p->ctx->stream, p->d_sd, p->d_planes_pipe, p->d_units_pipe + p->pipe_level_begin[lv],
                      p->pipe_level_count[lv], p->shape_pipe, buf + p->pipe_level_begin[lv]);
    if (rc) return rc;
  }
  if ((rc = launch_sum(p->ctx->stream, buf, n, buf + n))) return rc;
  KR_HIP_CHECK(hipMemcpyAsync(population, buf + n, sizeof(double), hipMemcpyDeviceToHost, p->ctx->stream));
  KR_HIP_CHECK(hipStreamSynchronize(p->ctx->stream));
  return KR_OK;
}

int kr_population(kr_problem *p, double *out) {
  if (!p || !out) return fail(KR_ERR_INVALID, "kr_population: null argument");
  {
    CallTimer t(p, "Population");
    int rc = launch_population(launch_ctx(p), p->d_pop_blocks, p->pop_nblocks, p->d_pop_partial, p->d_pop_out);
    if (rc) return rc;
  }
  KR_HIP_CHECK(hipMemcpyAsync(out, p->d_pop_out, sizeof(double), hipMemcpyDeviceToHost, p->ctx->stream));
  KR_HIP_CHECK(hipStreamSynchronize(p->ctx->stream));
  return KR_OK;
}

int kr_problem_set_timing(kr_problem *p, int32_t enable) {
  if (!p) return fail(KR_ERR_INVALID, "kr_problem_set_timing: null problem");
  p->timing = enable ? 1 : 0;
  return KR_OK;
}

int kr_problem_timer(kr_problem *p, const char *name, double *seconds, int64_t *count) {
  if (!p || !name) return fail(KR_ERR_INVALID, "kr_problem_timer: null argument");
  int rc = flush_timers(p);
  if (rc) return rc;
  auto it = p->timers.find(name);
  if (seconds) *seconds = it == p->timers.end() ? 0.0 : it->second.first;
  if (count) *count = it == p->timers.end() ? 0 : it->second.second;
  return KR_OK;
}

int kr_problem_timer_reset(kr_problem *p) {
  if (!p) return fail(KR_ERR_INVALID, "kr_problem_timer_reset: null problem");
  int rc = flush_timers(p);
  p->timers.clear();
  return rc;
}

}  // extern "C"
