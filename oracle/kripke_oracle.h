/*
 * ORACLE — test infrastructure only.
 *
 * CPU restatement of LLNL/Kripke's source-iteration path (problem generation,
 * the six Kernel:: bodies, SteadyStateSolver and the on-rank SweepSolver
 * queue).  It is the checker for the HIP product in kripke_amd/ and, built
 * with OpenMP, the CPU baseline that bench.py times beside the GPU.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 *
 * Two builds exist (oracle/Makefile):
 *   libkripke_oracle_ref.so  reference arithmetic: every `a += b*c` is a
 *                            rounded multiply then a rounded add
 *                            (-ffp-contract=off), as in the reference's
 *                            RAJA lambdas compiled for x86-64 without FMA.
 *   libkripke_oracle_fma.so  identical loops, but the three accumulations
 *                            that the HIP kernels fuse (LTimes, LPlusTimes,
 *                            sweep numerator) use fma(); the HIP path must
 *                            match this build bit for bit.
 *
 * Parity pins (what each function is checked against) are listed in
 * tests/test_oracle_pins.py; every function below is pinned there.
 */
#ifndef KRIPKE_ORACLE_H
#define KRIPKE_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Same field order and meaning as kr_input_vars in include/kripke_amd.h.
 * Mirrors InputVariables (src/Kripke/InputVariables.h:19-46) plus the
 * seeded-geometry knobs of the synthetic inputs. */
typedef struct ko_input {
  int32_t nx, ny, nz;
  int32_t num_directions;
  int32_t quad_num_polar, quad_num_azimuthal; /* 0,0 = dummy S2 */
  int32_t num_groups;
  int32_t legendre_order;
  int32_t layout;            /* 0..5 = DGZ DZG GDZ GZD ZDG ZGD */
  int32_t num_dirsets, num_groupsets;
  int32_t num_zonesets[3];
  int32_t num_material_subsamples;
  int32_t geometry;          /* 0 = reference Kobayashi-3i, 1 = seeded boxes in a 100 cm cube */
  uint64_t seed;
  double sigt[3], sigs[3];   /* per material; overwritten by the seeded generator */
  double source_strength;
  int32_t parallel_method;   /* 0 = sweep (SweepComm), 1 = Block Jacobi (BlockJacobiComm) */
} ko_input;

typedef struct ko_problem ko_problem;

ko_problem *ko_create(const ko_input *in);
void ko_destroy(ko_problem *p);
const char *ko_error(void);

/* Decomposition / sizes */
int32_t ko_num_sdoms(const ko_problem *p);
int32_t ko_num_pr(const ko_problem *p);
/* info[] = {gset, dset, zx, zy, zz, nd, ng, ni, nj, nk, dlo, glo, ilo, jlo, klo, pr, r,
 *           up0, up1, up2, down0, down1, down2, id, jd, kd} */
int32_t ko_sdom_info(const ko_problem *p, int32_t sdom, int64_t *info /* 26 */);
/* material cross sections actually used (after seeding) */
void ko_materials(const ko_problem *p, double *sigt3, double *sigs3);

/* Field access by reference name: "psi","rhs","i_plane","j_plane","k_plane"
 * (chunk = sdom), "phi","phi_out","sigt_zonal" (chunk = PR chunk id),
 * "ell","ell_plus","xcos","ycos","zcos","w","id","jd","kd" (chunk = dset),
 * "dx","dy","dz" (global per axis, chunk ignored), "volume" (chunk = R id),
 * "mixelem_to_zone","mixelem_to_material","mixelem_to_fraction",
 * "zone_to_mixelem","zone_to_num_mixelem" (chunk = R id),
 * "sigs" ([mat][n][g][gp], chunk ignored), "moment_to_legendre".
 * Returns element count; copies (as double) into dst when dst != NULL. */
int64_t ko_get(const ko_problem *p, const char *name, int32_t chunk, double *dst);
int64_t ko_set(ko_problem *p, const char *name, int32_t chunk, const double *src);

/* Kernel entry points (src/Kripke/Kernel.h:19-34).  omp != 0 selects the
 * OpenMP loop orders, omp == 0 the Sequential ones. */
void ko_kconst_all(ko_problem *p, const char *name, double value);
void ko_ltimes(ko_problem *p, int omp);
void ko_lplustimes(ko_problem *p, int omp);
void ko_scattering(ko_problem *p, int omp);
void ko_source(ko_problem *p, int omp);
void ko_sweep_subdomain(ko_problem *p, int32_t sdom, int omp);
void ko_sweep_solver(ko_problem *p, int omp);
void ko_sweep_solver_bj(ko_problem *p, int omp);
/* literal = sum in the Sequential loop order (or OpenMP reduction when omp);
 * compensated = Neumaier sum of the same rounded terms. */
void ko_population(ko_problem *p, int omp, double *literal, double *compensated);

/* One SteadyStateSolver iteration (src/Kripke/SteadyStateSolver.cpp:44-104).
 * Writes the population pair; phi stays readable through ko_get. */
void ko_iterate(ko_problem *p, int omp, double *literal, double *compensated);
/* Zero psi (SteadyStateSolver.cpp:39) */
void ko_init_psi(ko_problem *p);

/* Timers with the reference names (src/Kripke/Timing.h): seconds, calls. */
double ko_timer(const ko_problem *p, const char *name, int64_t *count);
void ko_timer_reset(ko_problem *p);
int32_t ko_omp_threads(void);

#ifdef __cplusplus
}
#endif
#endif
