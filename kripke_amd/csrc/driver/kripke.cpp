// The kripke command (src/kripke.cpp:37-516): parses the reference's options,
// prints the input summary, generates the problem (src/Kripke/Generate.cpp),
// runs SteadyStateSolver on the GPU through the library, prints the timers and
// the figures of merit.  One process drives one GPU: --procs must be 1,1,1.
// Options that exist only here: --mode, --geometry, --seed, --device.
#include <strings.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <sstream>
#include <string>
#include <vector>

#include "../host/kripke_host.h"

using namespace Kripke;
using namespace Kripke::Core;

namespace {

const char *LAYOUT_NAMES[6] = {"DGZ", "DZG", "GDZ", "GZD", "ZDG", "ZGD"};

// InputVariables defaults (src/Kripke/InputVariables.cpp:17-46) plus the
// driver-only settings.
struct DriverVars {
  kr_input_vars in{};
  int niter = 10;
  int npx = 1, npy = 1, npz = 1;
  int mode = 1;  // 0: the reference call sequence, 1: on-device pipeline
  int device = 0;
  bool layout_auto = false;  // --layout auto: kr_host_autotune_nesting picks it
  std::string run_name = "kripke";
  DriverVars() {
    in.nx = in.ny = in.nz = 16;
    in.num_directions = 96;
    in.quad_num_polar = in.quad_num_azimuthal = 0;
    in.num_groups = 32;
    in.legendre_order = 4;
    in.layout = KR_DGZ;
    in.num_dirsets = 8;
    in.num_groupsets = 2;
    in.num_zonesets[0] = in.num_zonesets[1] = in.num_zonesets[2] = 1;
    in.num_material_subsamples = 4;
    in.geometry = 0;  // Kobayashi-3i, the reference problem
    in.seed = 20261202;
    const double st[3] = {0.1, 0.0001, 0.1}, ss[3] = {0.05, 0.00005, 0.05};
    for (int m = 0; m < 3; ++m) {
      in.sigt[m] = st[m];
      in.sigs[m] = ss[m];
    }
    in.source_strength = 1.0;
    in.parallel_method = 0;
  }
};

[[noreturn]] void usage() {
  DriverVars def;
  printf("Usage:  kripke [options...]\n\n");
  printf("Problem Size Options:\n");
  printf("---------------------\n");
  printf("  --groups <ngroups>     Number of energy groups\n");
  printf("                         Default:  --groups %d\n\n", def.in.num_groups);
  printf("  --legendre <lorder>    Scattering Legendre Expansion Order (0, 1, ...)\n");
  printf("                         Default:  --legendre %d\n\n", def.in.legendre_order);
  printf("  --quad [<ndirs>|<polar>:<azim>]\n");
  printf("                         Dummy S2 with <ndirs> points, or Gauss-Legendre\n");
  printf("                         with <polar> by <azim> points\n");
  printf("                         Default:  --quad %d\n\n", def.in.num_directions);
  printf("  --zones <x,y,z>        Number of zones in x,y,z\n");
  printf("                         Default:  --zones %d,%d,%d\n\n", def.in.nx, def.in.ny, def.in.nz);
  printf("\nPhysics Parameters:\n");
  printf("-------------------\n");
  printf("  --sigt <st0,st1,st2>   Total material cross-sections\n");
  printf("                         Default:   --sigt %lf,%lf,%lf\n\n", def.in.sigt[0], def.in.sigt[1], def.in.sigt[2]);
  printf("  --sigs <ss0,ss1,ss2>   Scattering material cross-sections\n");
  printf("                         Default:   --sigs %lf,%lf,%lf\n\n", def.in.sigs[0], def.in.sigs[1], def.in.sigs[2]);
  printf("  --geometry <g>         kobayashi (the reference problem) or boxes\n");
  printf("                         (seeded material boxes in a 100 cm cube)\n");
  printf("                         Default:   --geometry kobayashi\n\n");
  printf("  --seed <n>             Seed of the boxes geometry\n");
  printf("                         Default:   --seed %llu\n\n", (unsigned long long)def.in.seed);
  printf("\nOn-Node Options:\n");
  printf("----------------\n");
  printf("  --arch <ARCH>          Architecture selection\n");
  printf("                         Available: HIP\n");
  printf("                         Default:   --arch HIP\n\n");
  printf("  --layout <LAYOUT>      Data layout and loop nesting order\n");
  printf("                         Available: DGZ,DZG,GDZ,GZD,ZDG,ZGD,auto\n");
  printf("                         (auto: the fastest, timed on the device)\n");
  printf("                         Default:   --layout DGZ\n\n");
  printf("  --device <n>           HIP device\n");
  printf("                         Default:   --device 0\n\n");
  printf("\nParallel Decomposition Options:\n");
  printf("-------------------------------\n");
  printf("  --procs <npx,npy,npz>  Number of ranks in each spatial dimension\n");
  printf("                         (one process per GPU: only 1,1,1)\n");
  printf("                         Default:  --procs 1,1,1\n\n");
  printf("  --dset <ds>            Number of direction-sets\n");
  printf("                         Must be a factor of 8, and divide evenly the number\n");
  printf("                         of quadrature points\n");
  printf("                         Default:  --dset %d\n\n", def.in.num_dirsets);
  printf("  --gset <gs>            Number of energy group-sets\n");
  printf("                         Must divide evenly the number energy groups\n");
  printf("                         Default:  --gset %d\n\n", def.in.num_groupsets);
  printf("  --zset <zx>,<zy>,<zz>  Number of zone-sets in x,y, and z\n");
  printf("                         Default:  --zset 1,1,1\n\n");
  printf("\nSolver Options:\n");
  printf("---------------\n");
  printf("  --niter <NITER>        Number of solver iterations to run\n");
  printf("                         Default:  --niter %d\n\n", def.niter);
  printf("  --pmethod <method>     Parallel solver method\n");
  printf("                         sweep: Full up-wind sweep (wavefront algorithm)\n");
  printf("                         bj: Block Jacobi\n");
  printf("                         Default: --pmethod sweep\n\n");
  printf("  --mode <m>             ref: the reference call sequence (fills, per-\n");
  printf("                         subdomain sweeps from a host queue)\n");
  printf("                         pipeline: fused fills, on-device sweep pipeline\n");
  printf("                         Default: --mode pipeline\n\n");
  exit(1);
}

struct CmdLine {
  std::vector<std::string> args;
  size_t cur = 0;
  CmdLine(int argc, char **argv) : args(argv + 1, argv + argc) {}
  bool atEnd() const { return cur >= args.size(); }
  std::string pop() {
    if (atEnd()) usage();
    return args[cur++];
  }
};

std::vector<std::string> split(const std::string &str, char delim) {
  std::vector<std::string> out;
  std::stringstream ss(str);
  std::string e;
  while (std::getline(ss, e, delim)) out.push_back(e);
  return out;
}

// InputVariables::checkValues (src/Kripke/InputVariables.cpp:52-109); true on error
bool check_values(const DriverVars &v) {
  const kr_input_vars &in = v.in;
  if (in.num_zonesets[0] <= 0 || in.num_zonesets[1] <= 0 || in.num_zonesets[2] <= 0) {
    printf("Number of zone-sets in each dim need to be greater than or equal to 1\n");
    return true;
  }
  if (in.num_groups < 1) {
    printf("Number of groups (%d) needs to be at least 1\n", in.num_groups);
    return true;
  }
  if (in.num_groupsets < 1 || in.num_groups % in.num_groupsets) {
    printf("Number of groups (%d) must be evenly divided by number of groupsets (%d)\n", in.num_groups,
           in.num_groupsets);
    return true;
  }
  if (in.num_directions < 8) {
    printf("Number of directions (%d) needs to be at least 8\n", in.num_directions);
    return true;
  }
  if (in.num_dirsets < 1 || (in.num_dirsets % 8 && in.num_dirsets < 8)) {
    printf("Number of direction sets (%d) must be a multiple of 8\n", in.num_dirsets);
    return true;
  }
  if (in.num_directions % in.num_dirsets) {
    printf("Number of directions (%d) must be evenly divided by number of directionsets(%d)\n", in.num_directions,
           in.num_dirsets);
    return true;
  }
  if (in.legendre_order < 0) {
    printf("Legendre scattering order (%d) must be >= 0\n", in.legendre_order);
    return true;
  }
  if (v.niter < 1) {
    printf("You must run at least one iteration (%d)\n", v.niter);
    return true;
  }
  if (v.npx != 1 || v.npy != 1 || v.npz != 1) {
    printf("One process drives one GPU here: --procs must be 1,1,1 (got %d,%d,%d)\n", v.npx, v.npy, v.npz);
    return true;
  }
  return false;
}

void parse(DriverVars &v, int argc, char **argv) {
  CmdLine cmd(argc, argv);
  while (!cmd.atEnd()) {
    const std::string opt = cmd.pop();
    if (opt == "-h" || opt == "--help") usage();
    else if (opt == "--name") v.run_name = cmd.pop();
    else if (opt == "--dset") v.in.num_dirsets = std::atoi(cmd.pop().c_str());
    else if (opt == "--gset") v.in.num_groupsets = std::atoi(cmd.pop().c_str());
    else if (opt == "--zset" || opt == "--zones" || opt == "--procs") {
      const auto p = split(cmd.pop(), ',');
      if (p.size() != 3) usage();
      int *dst = opt == "--zset" ? v.in.num_zonesets : nullptr;
      int a = std::atoi(p[0].c_str()), b = std::atoi(p[1].c_str()), c = std::atoi(p[2].c_str());
      if (dst) {
        dst[0] = a; dst[1] = b; dst[2] = c;
      } else if (opt == "--zones") {
        v.in.nx = a; v.in.ny = b; v.in.nz = c;
      } else {
        v.npx = a; v.npy = b; v.npz = c;
      }
    } else if (opt == "--pmethod") {
      const std::string m = cmd.pop();
      if (!strcasecmp(m.c_str(), "sweep")) v.in.parallel_method = 0;
      else if (!strcasecmp(m.c_str(), "bj")) v.in.parallel_method = 1;
      else usage();
    } else if (opt == "--groups") v.in.num_groups = std::atoi(cmd.pop().c_str());
    else if (opt == "--quad") {
      const auto p = split(cmd.pop(), ':');
      if (p.size() == 1) {
        v.in.num_directions = std::atoi(p[0].c_str());
        v.in.quad_num_polar = v.in.quad_num_azimuthal = 0;
      } else if (p.size() == 2) {
        v.in.quad_num_polar = std::atoi(p[0].c_str());
        v.in.quad_num_azimuthal = std::atoi(p[1].c_str());
        v.in.num_directions = v.in.quad_num_polar * v.in.quad_num_azimuthal;
      } else {
        usage();
      }
    } else if (opt == "--legendre") v.in.legendre_order = std::atoi(cmd.pop().c_str());
    else if (opt == "--sigs" || opt == "--sigt") {
      const auto p = split(cmd.pop(), ',');
      if (p.size() != 3) usage();
      double *dst = opt == "--sigs" ? v.in.sigs : v.in.sigt;
      for (int m = 0; m < 3; ++m) dst[m] = std::atof(p[m].c_str());
    } else if (opt == "--niter") v.niter = std::atoi(cmd.pop().c_str());
    else if (opt == "--arch") {
      const std::string a = cmd.pop();
      if (strcasecmp(a.c_str(), "HIP")) {
        printf("Architecture %s is not built here (available: HIP)\n", a.c_str());
        exit(1);
      }
    } else if (opt == "--layout") {
      const std::string l = cmd.pop();
      int found = -1;
      for (int i = 0; i < 6; ++i)
        if (!strcasecmp(l.c_str(), LAYOUT_NAMES[i])) found = i;
      v.layout_auto = !strcasecmp(l.c_str(), "auto");
      if (v.layout_auto) found = KR_DGZ;
      if (found < 0) {
        printf("Unknown layout %s\n", l.c_str());
        exit(1);
      }
      v.in.layout = found;
    } else if (opt == "--mode") {
      const std::string m = cmd.pop();
      if (m == "ref") v.mode = 0;
      else if (m == "pipeline") v.mode = 1;
      else usage();
    } else if (opt == "--geometry") {
      const std::string g = cmd.pop();
      if (g == "kobayashi") v.in.geometry = 0;
      else if (g == "boxes") v.in.geometry = 1;
      else usage();
    } else if (opt == "--seed") v.in.seed = std::strtoull(cmd.pop().c_str(), nullptr, 10);
    else if (opt == "--device") v.device = std::atoi(cmd.pop().c_str());
    else {
      printf("Unknown option %s\n", opt.c_str());
      usage();
    }
  }
}

void print_inputs(const DriverVars &v) {
  const kr_input_vars &in = v.in;
  printf("\nInput Parameters\n");
  printf("================\n\n");
  printf("  Problem Size:\n");
  printf("    Zones:                 %d x %d x %d  (%d total)\n", in.nx, in.ny, in.nz, in.nx * in.ny * in.nz);
  printf("    Groups:                %d\n", in.num_groups);
  printf("    Legendre Order:        %d\n", in.legendre_order);
  printf("    Quadrature Set:        ");
  if (in.quad_num_polar == 0) printf("Dummy S2 with %d points\n", in.num_directions);
  else
    printf("Gauss-Legendre, %d polar, %d azimuthal (%d points)\n", in.quad_num_polar, in.quad_num_azimuthal,
           in.num_directions);
  printf("\n  Physical Properties:\n");
  printf("    Total X-Sec:           sigt=[%lf, %lf, %lf]\n", in.sigt[0], in.sigt[1], in.sigt[2]);
  printf("    Scattering X-Sec:      sigs=[%lf, %lf, %lf]\n", in.sigs[0], in.sigs[1], in.sigs[2]);
  printf("    Geometry:              %s\n", in.geometry == 0 ? "Kobayashi-3i" : "seeded material boxes");
  printf("\n  Solver Options:\n");
  printf("    Number iterations:     %d\n", v.niter);
  printf("    Call sequence:         %s\n", v.mode == 0 ? "reference (ref)" : "on-device pipeline");
  printf("\n  MPI Decomposition Options:\n");
  printf("    Total MPI tasks:       1\n");
  printf("    Spatial decomp:        %d x %d x %d MPI tasks\n", v.npx, v.npy, v.npz);
  printf("    Block solve method:    %s\n", in.parallel_method == 1 ? "Block Jacobi" : "Sweep");
  printf("\n  Per-Task Options:\n");
  printf("    DirSets/Directions:    %d sets, %d directions/set\n", in.num_dirsets,
         in.num_directions / in.num_dirsets);
  printf("    GroupSet/Groups:       %d sets, %d groups/set\n", in.num_groupsets, in.num_groups / in.num_groupsets);
  printf("    Zone Sets:             %d x %d x %d\n", in.num_zonesets[0], in.num_zonesets[1], in.num_zonesets[2]);
  printf("    Architecture:          HIP (device %d)\n", v.device);
  printf("    Data Layout:           %s\n", v.layout_auto ? "auto (chosen by the autotuner)" : LAYOUT_NAMES[in.layout]);
}

// Memory breakdown of the double fields (src/Kripke/Generate.cpp:56-94)
void print_fields(const DataStore &ds) {
  std::vector<std::string> names;
  for (const auto &n : ds.getVariableList())
    if (ds.isVariableType<Field<double>>(n)) names.push_back(n);
  std::sort(names.begin(), names.end());
  printf("\n  Memory breakdown of Field variables:\n");
  printf("  Field Variable            Num Elements    Megabytes\n");
  printf("  --------------            ------------    ---------\n");
  unsigned long total = 0;
  for (const auto &n : names) {
    const unsigned long sz = (unsigned long)ds.getVariable<Field<double>>(n).totalElements();
    total += sz;
    printf("  %-24s  %12lu %12.3lf\n", n.c_str(), sz, (double)sz * 8.0 / 1024.0 / 1024.0);
  }
  printf("  --------                  ------------    ---------\n");
  printf("  TOTAL                     %12lu %12.3lf\n", total, (double)total * 8.0 / 1024.0 / 1024.0);
  printf("\n  Generation Complete!\n");
}

}  // namespace

int main(int argc, char **argv) {
  printf("\n");
  printf("  kripke: discrete-ordinates transport source iteration (LLNL Kripke\n");
  printf("  interface) on AMD Instinct MI355X\n\n");
  printf("Compilation Options:\n");
  printf("  Architecture:           HIP (gfx950)\n");
  printf("  CHAI Enabled:           No\n");
  printf("  CUDA Enabled:           No\n");
  printf("  MPI Enabled:            No\n");
  printf("  OpenMP Enabled:         No\n");
  printf("  Caliper Enabled:        No (roctx ranges per timer)\n");

  DriverVars v;
  parse(v, argc, argv);
  if (check_values(v)) return 1;
  print_inputs(v);

  kr_context *ctx = nullptr;
  if (kr_context_create(v.device, &ctx) != KR_OK) {
    printf("ERROR: %s\n", kr_last_error());
    return 1;
  }
  if (v.layout_auto) {
    int32_t nest = KR_DGZ;
    double ms = 0.0;
    if (kr_host_autotune_nesting(ctx, &v.in, &nest, &ms) != KR_OK) {
      printf("ERROR: %s\n", kr_last_error());
      kr_context_destroy(ctx);
      return 1;
    }
    v.in.layout = nest;
    printf("\nLayout chosen by the autotuner: %s (%.3f ms per source iteration)\n", LAYOUT_NAMES[nest], ms);
  }
  int rc = 0;
  try {
    DataStore ds;
    auto *timing = new Timing();
    timing->enabled = true;
    timing->sync = true;  // each timer waits for its device work, as the reference's synchronous calls do
    timing->ctx = ctx;
    ds.addVariable("timing", timing);
    printf("\nGenerating Problem\n");
    printf("==================\n");
    generateProblem(ds, v.in, ctx);
    print_fields(ds);

    SteadyStateSolver(ds, (size_t)v.niter, v.in.parallel_method == 1, v.mode);
    kr_synchronize(ctx);

    timing->print();

    // Figures of merit (src/kripke.cpp:484-505)
    const size_t num_unknowns =
        (size_t)v.in.num_groups * (size_t)v.in.num_directions * (size_t)v.in.nx * v.in.ny * v.in.nz;
    const size_t num_iter = timing->getCount("SweepSolver");
    const double solve_time = timing->getTotal("Solve");
    const double iter_time = solve_time / (double)num_iter;
    const double grind_time = iter_time / (double)num_unknowns;
    const double throughput = (double)num_unknowns / iter_time;
    // the on-device pipeline (--mode pipeline) sweeps whole zone-set stages
    // and never calls Kernel::sweepSubdomain, so there is no SweepSubdomain
    // time to put over SweepSolver's
    const bool have_sdom_timer = timing->getCount("SweepSubdomain") > 0;
    const double sweep_eff = 100.0 * timing->getTotal("SweepSubdomain") / timing->getTotal("SweepSolver");
    printf("\n");
    printf("Figures of Merit\n");
    printf("================\n\n");
    printf("  Throughput:         %e [unknowns/(second/iteration)]\n", throughput);
    printf("  Grind time :        %e [(seconds/iteration)/unknowns]\n", grind_time);
    if (have_sdom_timer)
      printf("  Sweep efficiency :  %4.5lf [100.0 * SweepSubdomain time / SweepSolver time]\n", sweep_eff);
    else
      printf("  Sweep efficiency :  n/a [no SweepSubdomain calls: the pipeline sweeps whole zone-set stages]\n");
    printf("  Number of unknowns: %lu\n", (unsigned long)num_unknowns);
  } catch (const std::exception &e) {
    printf("ERROR: %s\n", e.what());
    rc = 1;
  }
  kr_context_destroy(ctx);
  if (rc == 0) printf("\nEND\n");
  return rc;
}
