// Host mirror of the reference's hot-path interface: generateProblem, the
// Kripke::Kernel entry points, SteadyStateSolver and SweepSolver.  Same
// names and call contracts as src/Kripke/{Generate.h,Kernel.h,
// SteadyStateSolver.h,SweepSolver.h}; the bodies call the C-ABI.
#pragma once
#include <chrono>
#include <map>
#include <string>
#include <vector>

#include "core.h"

namespace Kripke {

using InputVariables = kr_input_vars;

// "al": ArchLayout (src/Kripke/ArchLayout.h).  One architecture here (HIP).
struct ArchLayout : public Core::BaseVar {
  int layout = KR_DGZ;
};

// Registered kernel problem ("kr_problem" in the DataStore).
struct KernelProblem : public Core::BaseVar {
  kr_problem *prob = nullptr;
  kr_context *ctx = nullptr;
  ~KernelProblem() override {
    if (prob) kr_problem_destroy(prob);
  }
};

// Material data kept after generation (sigt/sigs per material).
struct MaterialData : public Core::BaseVar {
  double sigt[3], sigs[3];
  double source_strength = 1.0;
};

// Named accumulating wall-clock timers with the reference names and
// semantics (src/Kripke/Timing.h:25-115, Timing.cpp): the count rises at
// start, the elapsed time accumulates at stop.  Off unless `enabled`.  With
// `sync` the stop first waits for the device, so a timer covers the work its
// calls enqueued (the reference's calls are synchronous).  Every timed block
// is also a roctx range of the same name (the reference's Caliper
// annotation, Timing.h:91-107).
struct Timing : public Core::BaseVar {
  struct T {
    double total = 0;
    int64_t count = 0;
  };
  bool enabled = false, sync = false;
  kr_context *ctx = nullptr;
  std::map<std::string, T> timers;
  T &get(const std::string &n) { return timers[n]; }
  double getTotal(const std::string &n) const {
    auto it = timers.find(n);
    return it == timers.end() ? 0.0 : it->second.total;
  }
  size_t getCount(const std::string &n) const {
    auto it = timers.find(n);
    return it == timers.end() ? 0 : (size_t)it->second.count;
  }
  // Timers table and the machine-readable lines (Timing.cpp:45-95)
  void print() const;
};

// KRIPKE_TIMER analogue: times the enclosing block when timing is enabled.
class BlockTimer {
 public:
  BlockTimer(Core::DataStore &ds, const char *name);
  ~BlockTimer();

 private:
  Timing *m_t = nullptr;
  const char *m_name;
  std::chrono::steady_clock::time_point m_t0;
};
#define KRIPKE_TIMER(DS, NAME) Kripke::BlockTimer kr_block_timer_##NAME(DS, #NAME)

void generateProblem(Core::DataStore &ds, const InputVariables &in, kr_context *ctx);

namespace Kernel {
void LTimes(Core::DataStore &ds);
void LPlusTimes(Core::DataStore &ds);
void scattering(Core::DataStore &ds);
void source(Core::DataStore &ds);
void sweepSubdomain(Core::DataStore &ds, SdomId sdom_id);
double population(Core::DataStore &ds);
// kConst over every chunk of a field / one chunk (src/Kripke/Kernel.h:37-55)
void kConst(Core::Field<double> &field, double value);
void kConst(Core::Field<double> &field, SdomId sdom_id, double value);
// Variants used by the on-device pipeline: start from +0.0 instead of
// reading a zero-filled field (same results, no fill pass).
void LTimesOverwrite(Core::DataStore &ds);
void LPlusTimesOverwrite(Core::DataStore &ds);
void scatteringOverwrite(Core::DataStore &ds);
}  // namespace Kernel

// src/Kripke/SweepSolver.cpp:23-87.  pipeline=true uses kr_sweep_solver
// (kr_sweep_solver_bj when block_jacobi).
void SweepSolver(Core::DataStore &ds, const std::vector<SdomId> &sdom_list, bool pipeline, bool block_jacobi = false);
// One iteration of src/Kripke/SteadyStateSolver.cpp:44-104.
double SteadyStateIteration(Core::DataStore &ds, int mode, bool block_jacobi = false);
// src/Kripke/SteadyStateSolver.cpp:24-115: psi = 0, max_iter iterations,
// the reference's per-iteration particle-count lines on stdout when
// `verbose`.  Returns 0.
int SteadyStateSolver(Core::DataStore &ds, size_t max_iter, bool block_jacobi, int mode, bool verbose = true);

// Thrown on any C-ABI failure inside the host mirror (KRIPKE_ABORT analogue).
struct KripkeError : public std::runtime_error {
  int code;
  KripkeError(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};
void check(int rc, const char *what);

}  // namespace Kripke
