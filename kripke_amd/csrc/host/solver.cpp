// Kripke::Kernel entry points, SweepSolver and the steady-state iteration
// of the host mirror, driving the C-ABI (src/Kripke/Kernel/*.cpp,
// src/Kripke/SweepSolver.cpp, src/Kripke/SteadyStateSolver.cpp).
#include <cstdio>
#include <cstring>

#include <rocprofiler-sdk-roctx/roctx.h>

#include "kripke_host.h"

using namespace Kripke;
using namespace Kripke::Core;

namespace {
kr_problem *kp(DataStore &ds) { return ds.getVariable<KernelProblem>("kr_problem").prob; }
}  // namespace

BlockTimer::BlockTimer(DataStore &ds, const char *name) : m_name(name) {
  if (!ds.isVariableDefined("timing")) return;
  Timing &t = ds.getVariable<Timing>("timing");
  if (!t.enabled) return;
  m_t = &t;
  roctxRangePushA(name);
  t.get(name).count++;
  m_t0 = std::chrono::steady_clock::now();
}

BlockTimer::~BlockTimer() {
  if (!m_t) return;
  if (m_t->sync && m_t->ctx) kr_synchronize(m_t->ctx);
  m_t->get(m_name).total += std::chrono::duration<double>(std::chrono::steady_clock::now() - m_t0).count();
  roctxRangePop();
}

void Timing::print() const {
  // std::map keeps the names sorted, as Timing.cpp:51-57 sorts them
  printf("\nTimers\n");
  printf("======\n\n");
  printf("  %-16s  %12s  %12s\n", "Timer", "Count", "Seconds");
  printf("  ----------------  ------------  ------------\n");
  for (const auto &kv : timers) printf("  %-16s  %12d  %12.5lf\n", kv.first.c_str(), (int)kv.second.count, kv.second.total);
  printf("\nTIMER_NAMES:");
  bool first = true;
  for (const auto &kv : timers) {
    printf("%s%s", first ? "" : ",", kv.first.c_str());
    first = false;
  }
  printf("\nTIMER_DATA:");
  first = true;
  for (const auto &kv : timers) {
    printf("%s%lf", first ? "" : ",", kv.second.total);
    first = false;
  }
  printf("\n");
}

void Kernel::kConst(Field<double> &field, SdomId sdom_id, double value) {
  check(kr_kconst(field.context(), field.device(sdom_id), (int64_t)field.size(sdom_id), value), "kConst");
}

void Kernel::kConst(Field<double> &field, double value) {
  for (SdomId s : field.getWorkList()) kConst(field, s, value);
}

void Kernel::LTimes(DataStore &ds) {
  KRIPKE_TIMER(ds, LTimes);
  check(kr_ltimes(kp(ds), 1), "LTimes");
}
void Kernel::LPlusTimes(DataStore &ds) {
  KRIPKE_TIMER(ds, LPlusTimes);
  check(kr_lplustimes(kp(ds), 1), "LPlusTimes");
}
void Kernel::scattering(DataStore &ds) {
  KRIPKE_TIMER(ds, Scattering);
  check(kr_scattering(kp(ds), 1), "scattering");
}
void Kernel::LTimesOverwrite(DataStore &ds) {
  KRIPKE_TIMER(ds, LTimes);
  check(kr_ltimes(kp(ds), 0), "LTimes");
}
void Kernel::LPlusTimesOverwrite(DataStore &ds) {
  KRIPKE_TIMER(ds, LPlusTimes);
  check(kr_lplustimes(kp(ds), 0), "LPlusTimes");
}
void Kernel::scatteringOverwrite(DataStore &ds) {
  KRIPKE_TIMER(ds, Scattering);
  check(kr_scattering(kp(ds), 0), "scattering");
}

void Kernel::source(DataStore &ds) {
  KRIPKE_TIMER(ds, Source);
  check(kr_source(kp(ds), ds.getVariable<MaterialData>("materials").source_strength), "source");
}

void Kernel::sweepSubdomain(DataStore &ds, SdomId sdom_id) {
  KRIPKE_TIMER(ds, SweepSubdomain);
  check(kr_sweep_subdomain(kp(ds), sdom_id), "sweepSubdomain");
}

double Kernel::population(DataStore &ds) {
  KRIPKE_TIMER(ds, Population);
  double part = 0.0;
  check(kr_population(kp(ds), &part), "population");
  return part;  // Comm::allReduceSumDouble is the identity on one rank
}

// SweepSolver (SweepSolver.cpp:23-87) with the on-rank SweepComm
// (ParallelComm.cpp:61-263) or, block_jacobi, BlockJacobiComm
// (BlockJacobiComm.cpp:23-116).  pipeline=false replays the reference queue
// on the host; pipeline=true hands the whole list to one device call.
void Kripke::SweepSolver(DataStore &ds, const std::vector<SdomId> &sdom_list, bool pipeline, bool block_jacobi) {
  KRIPKE_TIMER(ds, SweepSolver);
  if (pipeline) {
    if (block_jacobi)
      check(kr_sweep_solver_bj(kp(ds), sdom_list.data(), (int32_t)sdom_list.size()), "SweepSolver");
    else
      check(kr_sweep_solver(kp(ds), sdom_list.data(), (int32_t)sdom_list.size()), "SweepSolver");
    return;
  }
  auto &ipl = ds.getVariable<Field<double>>("i_plane");
  auto &jpl = ds.getVariable<Field<double>>("j_plane");
  auto &kpl = ds.getVariable<Field<double>>("k_plane");
  Field<double> *planes[3] = {&ipl, &jpl, &kpl};
  Kernel::kConst(ipl, 0.0);
  Kernel::kConst(jpl, 0.0);
  Kernel::kConst(kpl, 0.0);
  auto &up = ds.getVariable<Field<int32_t>>("upwind");
  if (block_jacobi) {
    // The communicator's fresh old planes are zero; the reversed kCopy in
    // addSubdomain and the on-rank sends of the first workRemaining() copy
    // them into every plane chunk, which are zero already.  Then every
    // subdomain is ready: FIFO, vacuum faces cleared, no sends on completion.
    for (SdomId s : sdom_list) {
      for (int dim = 0; dim < 3; ++dim)
        if (up.host(s)[dim] == -1) Kernel::kConst(*planes[dim], s, 0.0);
      Kernel::sweepSubdomain(ds, s);
    }
    return;
  }
  auto &down = ds.getVariable<Field<int32_t>>("downwind");
  std::vector<SdomId> q_ids;
  std::vector<int> q_dep;
  for (SdomId s : sdom_list) {  // postRecvs: count on-rank upwind dependencies
    int n = 0;
    for (int dim = 0; dim < 3; ++dim)
      if (up.host(s)[dim] >= 0) ++n;
    q_ids.push_back(s);
    q_dep.push_back(n);
  }
  while (!q_ids.empty()) {
    int pick = -1;
    for (size_t i = 0; i < q_ids.size(); ++i)
      if (q_dep[i] == 0) {
        pick = (int)i;
        break;
      }
    if (pick < 0) throw KripkeError(KR_ERR_INVALID, "SweepSolver: no ready subdomain (missing upwind subdomain)");
    const SdomId s = q_ids[pick];
    for (int dim = 0; dim < 3; ++dim)
      if (up.host(s)[dim] == -1) Kernel::kConst(*planes[dim], s, 0.0);
    Kernel::sweepSubdomain(ds, s);
    q_ids.erase(q_ids.begin() + pick);
    q_dep.erase(q_dep.begin() + pick);
    for (int dim = 0; dim < 3; ++dim) {  // postSends, on-rank branch
      const int dn = down.host(s)[dim];
      if (dn < 0) continue;
      for (size_t i = 0; i < q_ids.size(); ++i)
        if (q_ids[i] == dn) {
          q_dep[i]--;
          break;
        }
      check(kr_kcopy(planes[dim]->context(), planes[dim]->device(dn), planes[dim]->device(s),
                     (int64_t)planes[dim]->size(s)),
            "plane hand-off");
    }
  }
}

// SteadyStateSolver.cpp:44-104.  mode 0: the reference call sequence;
// mode 1: fills folded into the kernels, the on-device sweep pipeline, and
// the population summed while the sweep writes psi.
double Kripke::SteadyStateIteration(DataStore &ds, int mode, bool block_jacobi) {
  auto &pspace = ds.getVariable<PartitionSpace>("pspace");
  const int n = (int)pspace.getNumSubdomains(SPACE_PQR);
  std::vector<SdomId> all(n);
  for (int i = 0; i < n; ++i) all[i] = i;
  if (mode == 0) {
    Kernel::kConst(ds.getVariable<Field<double>>("phi"), 0.0);
    Kernel::LTimes(ds);
    Kernel::kConst(ds.getVariable<Field<double>>("phi_out"), 0.0);
    Kernel::scattering(ds);
    Kernel::source(ds);
    Kernel::kConst(ds.getVariable<Field<double>>("rhs"), 0.0);
    Kernel::LPlusTimes(ds);
    SweepSolver(ds, all, false, block_jacobi);
  } else {
    Kernel::LTimesOverwrite(ds);
    Kernel::scatteringOverwrite(ds);
    Kernel::source(ds);
    Kernel::LPlusTimesOverwrite(ds);
    if (block_jacobi) {
      SweepSolver(ds, all, true, true);
      return Kernel::population(ds);
    }
    // SweepSolver + population in one pass over psi.  psi is the
    // reference's; the plane fields are not (kr_sweep_solver writes faces
    // straight downwind, so only the domain's downwind-boundary chunks end
    // with their own outgoing faces), and nothing here reads them afterwards
    double part = 0.0;
    {
      KRIPKE_TIMER(ds, SweepSolver);
      check(kr_sweep_solver_population(kp(ds), all.data(), (int32_t)all.size(), &part), "SweepSolver");
    }
    return part;
  }
  return Kernel::population(ds);
}

int Kripke::SteadyStateSolver(DataStore &ds, size_t max_iter, bool block_jacobi, int mode, bool verbose) {
  KRIPKE_TIMER(ds, Solve);
  if (verbose) {
    printf("\n");
    printf("Steady State Solve\n");
    printf("==================\n\n");
  }
  Kernel::kConst(ds.getVariable<Field<double>>("psi"), 0.0);  // SteadyStateSolver.cpp:39
  double part_last = 0.0;
  for (size_t iter = 0; iter < max_iter; ++iter) {
    const double part = SteadyStateIteration(ds, mode, block_jacobi);
    if (verbose) {
      printf("  iter %d: particle count=%e, change=%e\n", (int)iter, part, (part - part_last) / part);
      fflush(stdout);
    }
    part_last = part;
  }
  if (verbose) printf("  Solver terminated\n");
  return 0;
}
