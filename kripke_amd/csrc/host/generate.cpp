// generateProblem for the host mirror: decomposition, energy groups,
// quadrature with L/L+ (real spherical harmonics), mesh and materials,
// cross sections and the solution fields (src/Kripke/Generate.cpp:24-94,
// src/Kripke/Generate/{Decomp,Energy,Quadrature,Space,Data}.cpp).
// The seeded box geometry of the benchmark configurations replaces the
// Kobayashi geometry when InputVariables::geometry == 1.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>

#include "kripke_host.h"

using namespace Kripke;
using namespace Kripke::Core;

namespace {

struct QuadraturePoint {  // Quadrature.cpp:23-36
  double xcos, ycos, zcos, w;
  int id, jd, kd, octant;
};

// Gauss-Legendre nodes and weights by Newton iteration on P_n
// (Quadrature.cpp:46-76).
void gauss_legendre(double x1, double x2, std::vector<double> &x, std::vector<double> &w, double eps) {
  const int n = (int)x.size(), m = (n + 1) / 2;
  const double xm = 0.5 * (x2 + x1), xl = 0.5 * (x2 - x1);
  for (int i = 1; i <= m; i++) {
    double z = cos(M_PI * (i - 0.25) / (n + 0.5)), z1, pp, p1, p2, p3;
    do {
      p1 = 1.0;
      p2 = 0.0;
      for (int j = 1; j <= n; j++) {
        p3 = p2;
        p2 = p1;
        p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
      }
      pp = n * (z * p1 - p2) / (z * z - 1.0);
      z1 = z;
      z = z1 - p1 / pp;
    } while (fabs(z - z1) > eps);
    x[i - 1] = xm - xl * z;
    x[n - i] = xm + xl * z;
    w[i - 1] = 2.0 * xl / ((1.0 - z * z) * pp * pp);
    w[n - i] = w[i - 1];
  }
}

double factorial(int n) {
  double f = 1.0;
  for (int i = n; i > 0; --i) f *= (double)i;
  return f;
}

// Associated Legendre P_n^m(x) with the Condon-Shortley phase, upward
// recurrence in n (Quadrature.cpp:93-138).
double assoc_legendre(int n, int m, double x) {
  double pmm = 1.0;
  if (m > 0) {
    const double somx2 = sqrt((1.0 - x) * (1.0 + x));
    double fact = 1.0;
    for (int i = 1; i <= m; i++) {
      pmm *= -fact * somx2;
      fact += 2.0;
    }
  }
  if (n == m) return pmm;
  double pmmp1 = x * (2 * m + 1) * pmm;
  if (n == m + 1) return pmmp1;
  double pnn = 0.0;
  for (int nn = m + 2; nn <= n; nn++) {
    pnn = (x * (2 * nn - 1) * pmmp1 - (nn + m - 1) * pmm) / (nn - m);
    pmm = pmmp1;
    pmmp1 = pnn;
  }
  return pnn;
}

// Real spherical harmonic Y_n^m at direction (mu, eta, xi)
// (Quadrature.cpp:140-183): azimuth from atan(|eta/(mu+1e-20)|) by quadrant.
double real_ynm(int n, int m, double mu, double eta, double xi) {
  const double tiny = 1.e-20;
  const double t = atan(fabs(eta / (mu + tiny)));
  double phi;
  if ((mu > 0) && (eta > 0)) phi = t;
  else if ((mu < 0) && (eta > 0)) phi = M_PI - fabs(t);
  else if ((mu < 0) && (eta < 0)) phi = M_PI + fabs(t);
  else phi = 2.0 * M_PI - fabs(t);
  const int am = std::abs(m);
  const double f1 = factorial(n - am), f2 = factorial(n + am);
  const double pnm = assoc_legendre(n, am, xi);
  const double arg = ((double)m) * phi;
  const double taum = (m >= 0) ? cos(arg) : sin(-arg);
  const double dm0 = (m == 0) ? 1.0 : 0.0;
  const double anm = sqrt(((2 * n + 1) * f1) / (2.0 * (1.0 + dm0) * M_PI * f2));
  return anm * pnm * taum;
}

bool by_octant_descending(QuadraturePoint const &a, QuadraturePoint const &b) { return b.octant < a.octant; }

// Quadrature.cpp:192-310: Gauss-Legendre polar x uniform azimuth product set
// sorted by octant, or the dummy S2 set with repeated directions.
std::vector<QuadraturePoint> quadrature_set(const InputVariables &in) {
  const int D = in.num_directions, per_oct = D / 8;
  std::vector<QuadraturePoint> dirs(D, QuadraturePoint{0, 0, 0, 0, 0, 0, 0, 0});
  const int np = in.quad_num_polar, na = in.quad_num_azimuthal;
  if (np > 0) {
    if (np % 4 || na % 2 || np * na != D)
      throw KripkeError(KR_ERR_INVALID, "Gauss-Legendre quadrature needs polar%4==0, azimuthal%2==0, polar*azimuthal==directions");
    std::vector<double> pc(np), pw(np), az(na), aw(na);
    gauss_legendre(-1.0, 1.0, pc, pw, DBL_EPSILON);
    const double dangle = 2.0 * M_PI / ((double)na);
    for (int i = 0; i < na; i++) {
      az[i] = (i == 0) ? dangle / 2.0 : az[i - 1] + dangle;
      aw[i] = dangle;
    }
    int d = 0;
    for (int i = 0; i < np; i++)
      for (int j = 0; j < na; j++) {
        const double xc = sqrt(1.0 - pc[i] * pc[i]) * cos(az[j]);
        const double yc = sqrt(1.0 - pc[i] * pc[i]) * sin(az[j]);
        const double zc = pc[i];
        QuadraturePoint &q = dirs[d++];
        q.w = pw[i] * aw[j];
        q.id = (xc > 0.) ? 1 : -1;
        q.jd = (yc > 0.) ? 1 : -1;
        q.kd = (zc > 0.) ? 1 : -1;
        q.octant = (q.id == -1 ? 1 : 0) + (q.jd == -1 ? 2 : 0) + (q.kd == -1 ? 4 : 0);
        q.xcos = std::abs(xc);
        q.ycos = std::abs(yc);
        q.zcos = std::abs(zc);
      }
    std::sort(dirs.begin(), dirs.end(), by_octant_descending);
  } else {
    const double mu = cos(M_PI / 4);
    const double eta = sqrt(1 - mu * mu) * cos(M_PI / 4);
    const double xi = sqrt(1 - mu * mu) * sin(M_PI / 4);
    int d = 0;
    for (int oct = 0; oct < 8; ++oct)
      for (int sd = 0; sd < per_oct; ++sd, ++d) {
        dirs[d].id = (oct & 1) ? 1 : -1;
        dirs[d].jd = (oct & 2) ? 1 : -1;
        dirs[d].kd = (oct & 4) ? 1 : -1;
        dirs[d].w = 4.0 * M_PI / (double)D;
        dirs[d].xcos = mu;
        dirs[d].ycos = eta;
        dirs[d].zcos = xi;
        dirs[d].octant = oct ^ 7;
      }
  }
  return dirs;
}

// Seeded geometry of the benchmark configurations: SplitMix64 stream;
// material-1 box then material-0 box (side U[10%,50%] of the edge, corner
// uniform), sigma_t = 10^U[-4,0], sigma_s = U[0.3,0.9]*sigma_t.
struct SplitMix64 {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
};

struct Box {
  double lo[3], hi[3];
  bool contains(double x, double y, double z) const {
    return x >= lo[0] && x <= hi[0] && y >= lo[1] && y <= hi[1] && z >= lo[2] && z <= hi[2];
  }
};

void seeded_geometry(uint64_t seed, double edge, Box boxes[2], double sigt[3], double sigs[3]) {
  SplitMix64 rng{seed};
  for (int b = 0; b < 2; ++b)
    for (int ax = 0; ax < 3; ++ax) {
      const double side = edge * (0.1 + 0.4 * rng.uniform());
      const double lo = (edge - side) * rng.uniform();
      boxes[b].lo[ax] = lo;
      boxes[b].hi[ax] = lo + side;
    }
  for (int m = 0; m < 3; ++m) sigt[m] = pow(10.0, -4.0 + 4.0 * rng.uniform());
  for (int m = 0; m < 3; ++m) sigs[m] = (0.3 + 0.6 * rng.uniform()) * sigt[m];
}

// Kobayashi-3i regions (Space.cpp:142-177).
int kobayashi(double x, double y, double z) {
  x = std::abs(x);
  y = std::abs(y);
  z = std::abs(z);
  if (x <= 10.0 && y <= 10.0 && z <= 10.0) return 0;
  if (x <= 10.0 && y <= 60.0 && z <= 10.0) return 1;
  if (x <= 40.0 && y >= 50.0 && y <= 60.0 && z <= 10.0) return 1;
  if (x >= 30.0 && x <= 40.0 && y >= 50.0 && y <= 60.0 && z <= 40.0) return 1;
  if (x >= 30.0 && x <= 40.0 && y >= 50.0 && z >= 30.0 && z <= 40.0) return 1;
  return 2;
}

void check_inputs(const InputVariables &in) {  // InputVariables.cpp:52-109
  auto bad = [](const char *m) { throw KripkeError(KR_ERR_INVALID, m); };
  if (in.num_zonesets[0] <= 0 || in.num_zonesets[1] <= 0 || in.num_zonesets[2] <= 0)
    bad("Number of zone-sets in each dim need to be greater than or equal to 1");
  if (in.num_groups < 1) bad("Number of groups needs to be at least 1");
  if (in.num_groupsets < 1 || in.num_groups % in.num_groupsets)
    bad("Number of groups must be evenly divided by number of groupsets");
  if (in.num_directions < 8) bad("Number of directions needs to be at least 8");
  if (in.num_dirsets < 1 || (in.num_dirsets % 8 && in.num_dirsets < 8))
    bad("Number of direction sets must be a multiple of 8");
  if (in.num_directions % in.num_dirsets) bad("Number of directions must be evenly divided by number of directionsets");
  if (in.legendre_order < 0) bad("Legendre scattering order must be >= 0");
  if (in.layout < 0 || in.layout > 5) bad("Unknown layout");
  if (in.nx < 1 || in.ny < 1 || in.nz < 1) bad("Zone counts must be positive");
  if (in.nx % in.num_zonesets[0] || in.ny % in.num_zonesets[1] || in.nz % in.num_zonesets[2])
    bad("Number of zones must evenly divide into the number of subdomains");  // Space.cpp:49,65,80
  if (in.num_material_subsamples < 1) bad("num_material_subsamples must be >= 1");
  if (in.geometry != 0 && in.geometry != 1) bad("geometry must be 0 (Kobayashi) or 1 (seeded boxes)");
  if (in.parallel_method != 0 && in.parallel_method != 1) bad("parallel_method must be 0 (sweep) or 1 (bj)");
}

}  // namespace

void Kripke::check(int rc, const char *what) {
  if (rc != KR_OK) throw KripkeError(rc, std::string(what) + ": " + kr_last_error());
}

void Kripke::generateProblem(DataStore &ds, const InputVariables &in, kr_context *ctx) {
  check_inputs(in);
  const bool dev = ctx != nullptr;
  const int L = in.layout;
  using I = Idx;
  if (!ds.isVariableDefined("timing")) ds.addVariable("timing", new Timing());  // a caller may preset it
  KRIPKE_TIMER(ds, Generate);
  ds.newVariable<ArchLayout>("al").layout = L;

  // Decomp: one rank, P=Q=1 ranks (Decomp.cpp:23-56)
  auto &ps = ds.newVariable<PartitionSpace>("pspace", in.num_groupsets, in.num_dirsets, in.num_zonesets[0],
                                            in.num_zonesets[1], in.num_zonesets[2]);

  // Energy (Energy.cpp:22-42)
  auto &set_group = ds.newVariable<RangeSet>("Set/Group", ps, SPACE_P,
                                             std::vector<size_t>(ps.getNumSubdomains(SPACE_P), in.num_groups / in.num_groupsets));
  ds.newVariable<GlobalRangeSet>("Set/GlobalGroup", ps, set_group.globalSize());

  // Quadrature (Quadrature.cpp:315-455)
  auto &set_dir = ds.newVariable<RangeSet>(
      "Set/Direction", ps, SPACE_Q, std::vector<size_t>(ps.getNumSubdomains(SPACE_Q), in.num_directions / in.num_dirsets));
  const int L1 = in.legendre_order + 1, M = L1 * L1;
  auto &set_moment = ds.newVariable<GlobalRangeSet>("Set/Moment", ps, M);
  auto &set_legendre = ds.newVariable<GlobalRangeSet>("Set/Legendre", ps, L1);
  auto *f_m2l = new Field<int64_t>(ctx, set_moment, {I::Moment}, L, true, dev);
  ds.addVariable("moment_to_legendre", f_m2l);
  {
    int nm = 0;
    for (int n = 0; n < L1; ++n)
      for (int m = -n; m <= n; ++m) f_m2l->host(0)[nm++] = n;
  }
  const auto quad = quadrature_set(in);
  auto mk_dir_d = [&](const char *nm) {
    auto *f = new Field<double>(ctx, set_dir, {I::Direction}, L, true, dev);
    ds.addVariable(nm, f);
    return f;
  };
  auto mk_dir_i = [&](const char *nm) {
    auto *f = new Field<int32_t>(ctx, set_dir, {I::Direction}, L, true, dev);
    ds.addVariable(nm, f);
    return f;
  };
  auto *f_xcos = mk_dir_d("quadrature/xcos"), *f_ycos = mk_dir_d("quadrature/ycos");
  auto *f_zcos = mk_dir_d("quadrature/zcos"), *f_w = mk_dir_d("quadrature/w");
  auto *f_id = mk_dir_i("quadrature/id"), *f_jd = mk_dir_i("quadrature/jd"), *f_kd = mk_dir_i("quadrature/kd");
  auto *f_oct = mk_dir_i("quadrature/octant");
  for (SdomId s : f_xcos->getWorkList()) {
    const int nd = (int)set_dir.size(s), dlo = (int)set_dir.lower(s);
    for (int d = 0; d < nd; ++d) {
      const QuadraturePoint &q = quad[d + dlo];
      f_xcos->host(s)[d] = q.xcos;
      f_ycos->host(s)[d] = q.ycos;
      f_zcos->host(s)[d] = q.zcos;
      f_w->host(s)[d] = q.w;
      f_id->host(s)[d] = q.id;
      f_jd->host(s)[d] = q.jd;
      f_kd->host(s)[d] = q.kd;
      f_oct->host(s)[d] = q.octant;
    }
  }
  auto &set_ell = ds.newVariable<ProductSet>("Set/Ell", ps, SPACE_Q, std::vector<const Set *>{&set_moment, &set_dir});
  auto &set_ellp =
      ds.newVariable<ProductSet>("Set/EllPlus", ps, SPACE_Q, std::vector<const Set *>{&set_dir, &set_moment});
  auto *f_ell = new Field<double>(ctx, set_ell, {I::Moment, I::Direction}, L, true, dev);
  ds.addVariable("ell", f_ell);
  auto *f_ellp = new Field<double>(ctx, set_ellp, {I::Direction, I::Moment}, L, true, dev);
  ds.addVariable("ell_plus", f_ellp);
  const double SQRT4PI = std::sqrt(4 * M_PI);
  for (SdomId s : f_ell->getWorkList()) {
    const int nd = (int)set_dir.size(s), dlo = (int)set_dir.lower(s);
    int nm = 0;
    for (int n = 0; n < L1; n++)
      for (int m = -n; m <= n; m++) {
        for (int d = 0; d < nd; ++d) {
          const QuadraturePoint &q = quad[d + dlo];
          const double xc = (q.id) * (q.xcos), yc = (q.jd) * (q.ycos), zc = (q.kd) * (q.zcos);
          const double ynm = real_ynm(n, m, xc, yc, zc);
          f_ell->at(s, {nm, d}) = q.w * ynm / SQRT4PI;
          f_ellp->at(s, {d, nm}) = ynm * SQRT4PI;
        }
        nm++;
      }
  }
  // subdomain adjacency (Quadrature.cpp:460-518); one rank, so global ids = local ids
  auto &set_dim = ds.newVariable<GlobalRangeSet>("Set/Dimension", ps, 3);
  auto &set_adj = ds.newVariable<ProductSet>("Set/Adjacency", ps, SPACE_PQR, std::vector<const Set *>{&set_dim});
  auto *f_up = new Field<int32_t>(ctx, set_adj, {I::Dimension}, L, true, false);
  ds.addVariable("upwind", f_up);
  auto *f_down = new Field<int32_t>(ctx, set_adj, {I::Dimension}, L, true, false);
  ds.addVariable("downwind", f_down);
  for (SdomId s : f_up->getWorkList()) {
    int c[5];
    ps.sdomIdToCoord(s, c);
    const int sweep[3] = {f_id->host(s)[0], f_jd->host(s)[0], f_kd->host(s)[0]};
    for (int dim = 0; dim < 3; ++dim) {
      int cu[5], cd[5];
      memcpy(cu, c, sizeof(c));
      memcpy(cd, c, sizeof(c));
      cu[2 + dim] -= sweep[dim];
      cd[2 + dim] += sweep[dim];
      const int n = ps.dim(2 + dim);
      f_up->host(s)[dim] = (cu[2 + dim] < 0 || cu[2 + dim] >= n) ? -1 : ps.coordToSdomId(cu);
      f_down->host(s)[dim] = (cd[2 + dim] < 0 || cd[2 + dim] >= n) ? -1 : ps.coordToSdomId(cd);
    }
  }

  // Space (Space.cpp:38-365)
  auto &set_zi = ds.newVariable<RangeSet>("Set/ZoneI", ps, SPACE_RX,
                                          std::vector<size_t>(ps.getNumSubdomains(SPACE_RX), in.nx / in.num_zonesets[0]));
  auto &set_zj = ds.newVariable<RangeSet>("Set/ZoneJ", ps, SPACE_RY,
                                          std::vector<size_t>(ps.getNumSubdomains(SPACE_RY), in.ny / in.num_zonesets[1]));
  auto &set_zk = ds.newVariable<RangeSet>("Set/ZoneK", ps, SPACE_RZ,
                                          std::vector<size_t>(ps.getNumSubdomains(SPACE_RZ), in.nz / in.num_zonesets[2]));
  auto &set_zone =
      ds.newVariable<ProductSet>("Set/Zone", ps, SPACE_R, std::vector<const Set *>{&set_zk, &set_zj, &set_zi});
  auto &set_zlin = ds.newVariable<ProductSet>("Set/ZoneLinear", ps, SPACE_R, std::vector<const Set *>{&set_zone});
  auto &set_mat = ds.newVariable<GlobalRangeSet>("Set/Material", ps, 3);
  auto &matdata = ds.newVariable<MaterialData>("materials");
  matdata.source_strength = in.source_strength;
  double xmin, xmax, ymin, ymax, zmin, zmax;
  Box boxes[2];
  if (in.geometry == 1) {
    xmin = ymin = zmin = 0.0;
    xmax = ymax = zmax = 100.0;
    seeded_geometry(in.seed, 100.0, boxes, matdata.sigt, matdata.sigs);
  } else {
    xmin = -60.0; xmax = 60.0; ymin = -100.0; ymax = 100.0; zmin = -60.0; zmax = 60.0;
    for (int m = 0; m < 3; ++m) {
      matdata.sigt[m] = in.sigt[m];
      matdata.sigs[m] = in.sigs[m];
    }
  }
  const double dx = (xmax - xmin) / set_zi.globalSize();
  const double dy = (ymax - ymin) / set_zj.globalSize();
  const double dz = (zmax - zmin) / set_zk.globalSize();
  auto *f_dx = new Field<double>(ctx, set_zi, {I::ZoneI}, L, true, dev);
  ds.addVariable("dx", f_dx);
  auto *f_dy = new Field<double>(ctx, set_zj, {I::ZoneJ}, L, true, dev);
  ds.addVariable("dy", f_dy);
  auto *f_dz = new Field<double>(ctx, set_zk, {I::ZoneK}, L, true, dev);
  ds.addVariable("dz", f_dz);
  for (SdomId s : f_dx->getWorkList()) std::fill(f_dx->host(s), f_dx->host(s) + f_dx->size(s), dx);
  for (SdomId s : f_dy->getWorkList()) std::fill(f_dy->host(s), f_dy->host(s) + f_dy->size(s), dy);
  for (SdomId s : f_dz->getWorkList()) std::fill(f_dz->host(s), f_dz->host(s) + f_dz->size(s), dz);
  const double zone_volume = dx * dy * dz;
  auto *f_vol = new Field<double>(ctx, set_zlin, {I::Zone}, L, true, dev);
  ds.addVariable("volume", f_vol);
  for (SdomId s : f_vol->getWorkList()) std::fill(f_vol->host(s), f_vol->host(s) + f_vol->size(s), zone_volume);

  const int ns = in.num_material_subsamples;
  const double sample_vol_frac = 1.0 / (double)(ns * ns * ns);
  std::vector<std::vector<double>> mix;  // per R chunk: 3 fractions per zone
  std::vector<size_t> num_mixed;
  for (SdomId s : set_zone.getWorkList()) {
    const double x0 = xmin + dx * set_zi.lower(s), y0 = ymin + dy * set_zj.lower(s), z0 = zmin + dz * set_zk.lower(s);
    const int ni = (int)set_zi.size(s), nj = (int)set_zj.size(s), nk = (int)set_zk.size(s);
    std::vector<double> fr(3 * (size_t)set_zone.size(s), 0.0);
    for (int k = 0; k < nk; k++)
      for (int j = 0; j < nj; j++)
        for (int i = 0; i < ni; i++) {
          const size_t zone = ((size_t)k * nj + j) * ni + i;
          const double xi = x0 + dx * i, yi = y0 + dy * j, zi = z0 + dz * k;
          for (int si = 0; si < ns; ++si)
            for (int sj = 0; sj < ns; ++sj)
              for (int sk = 0; sk < ns; ++sk) {
                const double x = xi + dx * (si + 1) / (ns + 1);
                const double y = yi + dy * (sj + 1) / (ns + 1);
                const double z = zi + dz * (sk + 1) / (ns + 1);
                int mat;
                if (in.geometry == 1) mat = boxes[1].contains(x, y, z) ? 0 : (boxes[0].contains(x, y, z) ? 1 : 2);
                else mat = kobayashi(x, y, z);
                fr[3 * zone + mat] += sample_vol_frac;
              }
        }
    size_t n = 0;
    for (size_t z = 0; z < set_zone.size(s); ++z)
      for (int m = 0; m < 3; ++m) n += fr[3 * z + m] > 0.0 ? 1 : 0;
    mix.push_back(std::move(fr));
    num_mixed.push_back(n);
  }
  auto &set_mix = ds.newVariable<RangeSet>("Set/MixElem", ps, SPACE_R, num_mixed);
  auto *f_m2z = new Field<int64_t>(ctx, set_mix, {I::MixElem}, L, true, dev);
  ds.addVariable("mixelem_to_zone", f_m2z);
  auto *f_m2m = new Field<int64_t>(ctx, set_mix, {I::MixElem}, L, true, dev);
  ds.addVariable("mixelem_to_material", f_m2m);
  auto *f_m2f = new Field<double>(ctx, set_mix, {I::MixElem}, L, true, dev);
  ds.addVariable("mixelem_to_fraction", f_m2f);
  auto *f_z2n = new Field<int32_t>(ctx, set_zlin, {I::Zone}, L, true, dev);
  ds.addVariable("zone_to_num_mixelem", f_z2n);
  auto *f_z2m = new Field<int64_t>(ctx, set_zlin, {I::Zone}, L, true, dev);
  ds.addVariable("zone_to_mixelem", f_z2m);
  {
    const auto &wl = set_zone.getWorkList();
    for (size_t c = 0; c < wl.size(); ++c) {
      SdomId s = wl[c];
      int64_t me = 0;
      for (size_t z = 0; z < set_zone.size(s); ++z) {
        f_z2m->host(s)[z] = me;
        int cnt = 0;
        double zone_frac = 0.0;
        for (int m = 0; m < 3; ++m) {
          const double f = mix[c][3 * z + m];
          if (f > 0.0) {
            f_m2z->host(s)[me] = (int64_t)z;
            f_m2m->host(s)[me] = m;
            f_m2f->host(s)[me] = f;
            zone_frac += f;
            ++cnt;
            ++me;
          }
        }
        if (zone_frac != 1.0) throw KripkeError(KR_ERR_INVALID, "Zone fraction wrong");  // Space.cpp:312
        f_z2n->host(s)[z] = cnt;
      }
      if ((size_t)me != set_mix.size(s)) throw KripkeError(KR_ERR_INVALID, "Mismatch in mixture info");
    }
  }
  auto &set_sigt = ds.newVariable<ProductSet>("Set/SigmaTZonal", ps, SPACE_PR,
                                              std::vector<const Set *>{&set_group, &set_zone});
  auto *f_sigt = new Field<double>(ctx, set_sigt, {I::Group, I::Zone}, L, true, dev);
  ds.addVariable("sigt_zonal", f_sigt);
  for (SdomId s : f_sigt->getWorkList()) {
    const int ng = (int)set_group.size(s);
    const int64_t nmix = (int64_t)set_mix.size(s);
    for (int g = 0; g < ng; ++g)
      for (int64_t m = 0; m < nmix; ++m) {
        const int64_t z = f_m2z->host(s)[m];
        f_sigt->at(s, {g, z}) += f_m2f->host(s)[m] * matdata.sigt[f_m2m->host(s)[m]];
      }
  }

  // Data (Data.cpp:22-106)
  auto &set_flux = ds.newVariable<ProductSet>("Set/Flux", ps, SPACE_PQR,
                                              std::vector<const Set *>{&set_dir, &set_group, &set_zone});
  auto &set_fm = ds.newVariable<ProductSet>("Set/FluxMoment", ps, SPACE_PR,
                                            std::vector<const Set *>{&set_moment, &set_group, &set_zone});
  auto &set_ip = ds.newVariable<ProductSet>("Set/IPlane", ps, SPACE_PQR,
                                            std::vector<const Set *>{&set_dir, &set_group, &set_zj, &set_zk});
  auto &set_jp = ds.newVariable<ProductSet>("Set/JPlane", ps, SPACE_PQR,
                                            std::vector<const Set *>{&set_dir, &set_group, &set_zi, &set_zk});
  auto &set_kp = ds.newVariable<ProductSet>("Set/KPlane", ps, SPACE_PQR,
                                            std::vector<const Set *>{&set_dir, &set_group, &set_zi, &set_zj});
  ds.addVariable("psi", new Field<double>(ctx, set_flux, {I::Direction, I::Group, I::Zone}, L, false, dev));
  ds.addVariable("rhs", new Field<double>(ctx, set_flux, {I::Direction, I::Group, I::Zone}, L, false, dev));
  ds.addVariable("phi", new Field<double>(ctx, set_fm, {I::Moment, I::Group, I::Zone}, L, false, dev));
  ds.addVariable("phi_out", new Field<double>(ctx, set_fm, {I::Moment, I::Group, I::Zone}, L, false, dev));
  ds.addVariable("i_plane", new Field<double>(ctx, set_ip, {I::Direction, I::Group, I::ZoneJ, I::ZoneK}, L, false, dev));
  ds.addVariable("j_plane", new Field<double>(ctx, set_jp, {I::Direction, I::Group, I::ZoneI, I::ZoneK}, L, false, dev));
  ds.addVariable("k_plane", new Field<double>(ctx, set_kp, {I::Direction, I::Group, I::ZoneI, I::ZoneJ}, L, false, dev));
  auto &set_gg = ds.getVariable<GlobalRangeSet>("Set/GlobalGroup");
  auto &set_sigs = ds.newVariable<ProductSet>("Set/SigmaS", ps, SPACE_NULL,
                                              std::vector<const Set *>{&set_mat, &set_legendre, &set_gg, &set_gg});
  auto *f_sigs =
      new Field<double>(ctx, set_sigs, {I::Material, I::Legendre, I::GlobalGroup, I::GlobalGroup}, L, true, dev);
  ds.addVariable("data/sigs", f_sigs);
  for (int mat = 0; mat < 3; ++mat)
    for (int64_t g = 0; g < (int64_t)set_gg.globalSize(); ++g) f_sigs->at(0, {mat, 0, g, g}) = matdata.sigs[mat];

  if (!dev) return;

  // upload setup fields and register the kernel problem
  for (const auto &name : ds.getVariableList())
    if (ds.isVariableType<FieldBase>(name)) check(ds.getVariable<FieldBase>(name).upload(), "field upload");
  const size_t nsdom = ps.getNumSubdomains(SPACE_PQR), npr = ps.getNumSubdomains(SPACE_PR),
               nr = ps.getNumSubdomains(SPACE_R);
  auto &psi = ds.getVariable<Field<double>>("psi");
  auto &rhs = ds.getVariable<Field<double>>("rhs");
  auto &ipl = ds.getVariable<Field<double>>("i_plane");
  auto &jpl = ds.getVariable<Field<double>>("j_plane");
  auto &kpl = ds.getVariable<Field<double>>("k_plane");
  auto &phi = ds.getVariable<Field<double>>("phi");
  auto &phio = ds.getVariable<Field<double>>("phi_out");
  std::vector<kr_sdom_desc> sd(nsdom);
  std::vector<kr_pr_desc> pr(npr);
  std::vector<kr_r_desc> rr(nr);
  for (size_t s = 0; s < nsdom; ++s) {
    kr_sdom_desc &S = sd[s];
    memset(&S, 0, sizeof(S));
    int c[5];
    ps.sdomIdToCoord((SdomId)s, c);
    S.gset = c[0]; S.dset = c[1]; S.zx = c[2]; S.zy = c[3]; S.zz = c[4];
    S.pr_chunk = (int32_t)ps.subdomainToSpace(SPACE_PR, (SdomId)s);
    S.r_chunk = (int32_t)ps.subdomainToSpace(SPACE_R, (SdomId)s);
    for (int dim = 0; dim < 3; ++dim) {
      S.upwind[dim] = f_up->host((SdomId)s)[dim];
      S.downwind[dim] = f_down->host((SdomId)s)[dim];
    }
    S.num_directions = (int32_t)set_dir.size((SdomId)s);
    S.num_groups = (int32_t)set_group.size((SdomId)s);
    S.ni = (int32_t)set_zi.size((SdomId)s);
    S.nj = (int32_t)set_zj.size((SdomId)s);
    S.nk = (int32_t)set_zk.size((SdomId)s);
    S.direction_lower = (int32_t)set_dir.lower((SdomId)s);
    S.group_lower = (int32_t)set_group.lower((SdomId)s);
    S.psi = psi.device((SdomId)s);
    S.rhs = rhs.device((SdomId)s);
    S.i_plane = ipl.device((SdomId)s);
    S.j_plane = jpl.device((SdomId)s);
    S.k_plane = kpl.device((SdomId)s);
    S.xcos = f_xcos->device((SdomId)s);
    S.ycos = f_ycos->device((SdomId)s);
    S.zcos = f_zcos->device((SdomId)s);
    S.w = f_w->device((SdomId)s);
    S.id = f_id->device((SdomId)s);
    S.jd = f_jd->device((SdomId)s);
    S.kd = f_kd->device((SdomId)s);
    S.ell = f_ell->device((SdomId)s);
    S.ell_plus = f_ellp->device((SdomId)s);
    S.dx = f_dx->device((SdomId)s);
    S.dy = f_dy->device((SdomId)s);
    S.dz = f_dz->device((SdomId)s);
  }
  for (size_t c = 0; c < npr; ++c) {
    SdomId s = ps.spaceToSubdomain(SPACE_PR, c);
    pr[c].phi = phi.device(s);
    pr[c].phi_out = phio.device(s);
    pr[c].sigt_zonal = f_sigt->device(s);
  }
  for (size_t c = 0; c < nr; ++c) {
    SdomId s = ps.spaceToSubdomain(SPACE_R, c);
    rr[c].volume = f_vol->device(s);
    rr[c].zone_to_mixelem = f_z2m->device(s);
    rr[c].zone_to_num_mixelem = f_z2n->device(s);
    rr[c].mixelem_to_zone = f_m2z->device(s);
    rr[c].mixelem_to_material = f_m2m->device(s);
    rr[c].mixelem_to_fraction = f_m2f->device(s);
    rr[c].num_mixelem = (int64_t)set_mix.size(s);
  }
  kr_problem_desc D{};
  D.nesting = L;
  D.num_sdoms = (int32_t)nsdom;
  D.num_pr = (int32_t)npr;
  D.num_r = (int32_t)nr;
  D.num_moments = M;
  D.num_legendre = L1;
  D.num_materials = 3;
  D.global_num_groups = (int32_t)set_gg.globalSize();
  D.sdoms = sd.data();
  D.pr = pr.data();
  D.r = rr.data();
  D.sigs = f_sigs->device(0);
  D.moment_to_legendre = f_m2l->device(0);
  auto *kp = new KernelProblem();
  kp->ctx = ctx;
  ds.addVariable("kr_problem", kp);
  check(kr_problem_create(ctx, &D, &kp->prob), "kr_problem_create");
}
