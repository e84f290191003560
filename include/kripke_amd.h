/*
 * kripke_amd.h — C-ABI of the MI355X (gfx950) source-iteration hot path.
 *
 * Drop-in boundary for LLNL/Kripke's Kripke::Kernel entry points
 * (src/Kripke/Kernel.h:19-34) and its sweep driver
 * (src/Kripke/SweepSolver.cpp:23-87).  The reference's host code keeps its
 * DataStore/Set/Field objects; each Kernel:: function looks its fields up by
 * name as it does today and hands raw device pointers, extents and the
 * nesting to the functions below.  Only plain C types cross this boundary.
 *
 * Conventions
 *  - Every function returns KR_OK (0) or a KR_ERR_* code; kr_last_error()
 *    gives a message for the calling thread.  Nothing throws across it.
 *  - All work is ordered on the context's HIP stream and is asynchronous,
 *    except kr_population (returns a scalar) and the explicit copies.
 *  - Field chunks are laid out in the DefaultOrder of the nesting
 *    (src/Kripke/VarTypes.h:68-101): the index type listed first has the
 *    longest stride.  Flat offsets are 64-bit.
 *  - A context and the problems on it are not thread-safe.
 */
#ifndef KRIPKE_AMD_H
#define KRIPKE_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KR_OK 0
#define KR_ERR_INVALID 1     /* bad argument or descriptor */
#define KR_ERR_HIP 2         /* HIP runtime error */
#define KR_ERR_NOMEM 3       /* device allocation failed */
#define KR_ERR_UNSUPPORTED 4 /* shape outside what the kernels were built for */
#define KR_ERR_NO_DEVICE 5   /* no HIP device available */

/* LayoutV (src/Kripke/ArchLayout.h:92-101), same numbering. */
typedef enum kr_nesting { KR_DGZ = 0, KR_DZG = 1, KR_GDZ = 2, KR_GZD = 3, KR_ZDG = 4, KR_ZGD = 5 } kr_nesting;

typedef struct kr_context kr_context;
typedef struct kr_problem kr_problem;

/* One entry per subdomain (SdomId), in SdomId order.  Pointers are device
 * pointers to that subdomain's chunk of the named reference field. */
typedef struct kr_sdom_desc {
  int32_t gset, dset, zx, zy, zz;    /* PartitionSpace coordinates (PartitionSpace.cpp:97) */
  int32_t pr_chunk, r_chunk;         /* chunk ids in SPACE_PR and SPACE_R */
  int32_t upwind[3], downwind[3];    /* "upwind"/"downwind" (Field_Adjacency), -1 = boundary */
  int32_t num_directions, num_groups;/* Set/Direction, Set/Group sizes */
  int32_t ni, nj, nk;                /* Set/ZoneI, ZoneJ, ZoneK sizes */
  int32_t direction_lower, group_lower; /* Set::lower of Direction and Group */
  int32_t reserved;
  double *psi, *rhs;                 /* "psi", "rhs" (Field_Flux) */
  double *i_plane, *j_plane, *k_plane; /* "i_plane", "j_plane", "k_plane" */
  const double *xcos, *ycos, *zcos, *w; /* "quadrature/{xcos,ycos,zcos,w}" */
  const int32_t *id, *jd, *kd;       /* "quadrature/{id,jd,kd}" */
  const double *ell, *ell_plus;      /* "ell" (Field_Ell), "ell_plus" (Field_EllPlus) */
  const double *dx, *dy, *dz;        /* "dx", "dy", "dz" */
} kr_sdom_desc;

/* One entry per SPACE_PR chunk (group set x zone set). */
typedef struct kr_pr_desc {
  double *phi, *phi_out;             /* "phi", "phi_out" (Field_Moments) */
  const double *sigt_zonal;          /* "sigt_zonal" (Field_SigmaTZonal) */
} kr_pr_desc;

/* One entry per SPACE_R chunk (zone set). */
typedef struct kr_r_desc {
  const double *volume;              /* "volume" (Field_Zone2Double) */
  const int64_t *zone_to_mixelem;    /* "zone_to_mixelem" (Field_Zone2MixElem) */
  const int32_t *zone_to_num_mixelem;/* "zone_to_num_mixelem" (Field_Zone2Int) */
  const int64_t *mixelem_to_zone;    /* "mixelem_to_zone" */
  const int64_t *mixelem_to_material;/* "mixelem_to_material" */
  const double *mixelem_to_fraction; /* "mixelem_to_fraction" */
  int64_t num_mixelem;               /* Set/MixElem size */
} kr_r_desc;

typedef struct kr_problem_desc {
  int32_t nesting;                   /* kr_nesting */
  int32_t num_sdoms, num_pr, num_r;
  int32_t num_moments;               /* Set/Moment size, (L+1)^2 */
  int32_t num_legendre;              /* Set/Legendre size, L+1 */
  int32_t num_materials;             /* Set/Material size */
  int32_t global_num_groups;         /* Set/GlobalGroup size */
  const kr_sdom_desc *sdoms;         /* host array, num_sdoms entries */
  const kr_pr_desc *pr;              /* host array, num_pr entries */
  const kr_r_desc *r;                /* host array, num_r entries */
  const double *sigs;                /* "data/sigs" (Field_SigmaS), device */
  const int64_t *moment_to_legendre; /* "moment_to_legendre", device */
} kr_problem_desc;

/* ---- context and device memory ---------------------------------------- */
int kr_context_create(int32_t device, kr_context **out);
int kr_context_destroy(kr_context *ctx);
const char *kr_last_error(void);
int kr_device_count(int32_t *count);
int kr_malloc(kr_context *ctx, size_t bytes, void **dptr);
int kr_free(kr_context *ctx, void *dptr);
/* Copies are ordered on the context stream and complete before returning. */
int kr_memcpy_htod(kr_context *ctx, void *dst, const void *src, size_t bytes);
int kr_memcpy_dtoh(kr_context *ctx, void *dst, const void *src, size_t bytes);
int kr_synchronize(kr_context *ctx);

/* ---- problem registration ---------------------------------------------- */
/* Copies the descriptor tables (not the field data) and precomputes the
 * sweep coefficients 2*cos/delta per (direction, zone index) exactly as
 * SweepSubdomain.cpp:79-81 computes them.  Field pointers must stay valid
 * until kr_problem_destroy. */
int kr_problem_create(kr_context *ctx, const kr_problem_desc *desc, kr_problem **out);
int kr_problem_destroy(kr_problem *prob);

/* ---- Kernel:: entry points --------------------------------------------- */
/* Kernel::kConst (src/Kripke/Kernel.h:37-55): chunk[i] = value. */
int kr_kconst(kr_context *ctx, double *chunk, int64_t num_elem, double value);
/* Element copy dst[i] = src[i]; the on-rank plane hand-off of
 * ParallelComm::postSends (src/Kripke/ParallelComm.cpp:157-161). */
int kr_kcopy(kr_context *ctx, double *dst, const double *src, int64_t num_elem);

/* Kernel::LTimes (src/Kripke/Kernel/LTimes.cpp:79-105):
 * phi(nm,g,z) (+)= sum over subdomains, directions of ell(nm,d)*psi(d,g,z).
 * accumulate != 0 adds to phi as the reference does; accumulate == 0
 * starts from +0.0, which equals kConst(phi,0) followed by LTimes. */
int kr_ltimes(kr_problem *prob, int32_t accumulate);
/* Kernel::LPlusTimes (src/Kripke/Kernel/LPlusTimes.cpp:68-92):
 * rhs(d,g,z) (+)= sum_nm ell_plus(d,nm)*phi_out(nm,g,z). */
int kr_lplustimes(kr_problem *prob, int32_t accumulate);
/* Kernel::scattering (src/Kripke/Kernel/Scattering.cpp:112-164):
 * phi_out(nm,g,z) (+)= sum_gp sigs_z(n,g,gp)*phi(nm,gp,z) over subdomain
 * pairs sharing a zone set. */
int kr_scattering(kr_problem *prob, int32_t accumulate);
/* Kernel::source (src/Kripke/Kernel/Source.cpp:83-115):
 * phi_out(0,g,z) += strength*fraction for material-0 mixed elements. */
int kr_source(kr_problem *prob, double source_strength);
/* Kernel::sweepSubdomain (src/Kripke/Kernel/SweepSubdomain.cpp:115-124):
 * diamond-difference sweep of one subdomain; reads the incoming faces from
 * its own plane chunks and leaves the outgoing faces there. */
int kr_sweep_subdomain(kr_problem *prob, int32_t sdom_id);
/* SweepSolver (src/Kripke/SweepSolver.cpp:23-87) with the on-rank
 * SweepComm hand-off: sweeps every listed subdomain in dependency order,
 * vacuum faces read as zero, outgoing faces written straight into the
 * downwind subdomain's plane chunk.  Upwind subdomains must be listed.
 * psi is the reference's.  The plane fields end differently: the reference
 * leaves every subdomain's outgoing faces in its own chunk (and copies them
 * downwind first, ParallelComm.cpp:156-162).  Here subdomain t's chunk along
 * an axis ends holding t's outgoing faces where t is on the domain's
 * downwind boundary on that axis (as in the reference), its upwind
 * neighbour's outgoing faces (t's incoming faces) where t has neighbours on
 * both sides, and is not written where t has a downwind neighbour only.  The
 * reference's exact plane state comes from kr_sweep_subdomain plus kr_kcopy
 * hand-offs in the queue order. */
int kr_sweep_solver(kr_problem *prob, const int32_t *sdom_list, int32_t num);
/* SweepSolver with block_jacobi = true (--pmethod bj): SweepSolver.cpp:23-87
 * with BlockJacobiComm (src/Kripke/ParallelComm/BlockJacobiComm.cpp:23-116)
 * on one rank.  The reference's communicator allocates fresh old_*_plane
 * fields (BlockJacobiComm.cpp:31-33; new double[], Field.h:68, so their
 * contents are indeterminate there) and copies them into the live planes
 * (the reversed kCopy, Kernel.h:63-80) and into each downwind subdomain
 * (ParallelComm.cpp:137-162) before any sweep.  Here the old planes are zero,
 * so every listed subdomain is swept with zero incoming faces, independently
 * of the others (one launch), and leaves its outgoing faces in its own plane
 * chunks; plane chunks of unlisted subdomains end zero.  Ids must be
 * distinct. */
int kr_sweep_solver_bj(kr_problem *prob, const int32_t *sdom_list, int32_t num);
/* SweepSolver followed by Kernel::population, the two calls the reference
 * makes back to back (src/Kripke/SteadyStateSolver.cpp:92-99).  The list
 * must name every subdomain once.  The population is summed by the sweep
 * kernels as they write psi (no second pass over psi).  Same psi as
 * kr_sweep_solver; the population agrees with kr_population to rounding
 * (deterministic, fixed-order sums).  Blocks until the value is on the host. */
int kr_sweep_solver_population(kr_problem *prob, const int32_t *sdom_list, int32_t num, double *population);
/* No counterpart in the reference; for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc).  kr_sweep_solver_population with the fused population's
 * per-work-unit partials and their sum in a static device array of
 * KR_NOALLOC_SCRATCH_DOUBLES doubles (1 MiB, held by each device the
 * library runs on until the process ends) instead of buffers the problem
 * allocates on first use: the call allocates nothing and records no timers.
 * Same psi and population as kr_sweep_solver_population, bit for bit.
 * Problems whose pipelined sweep has KR_NOALLOC_SCRATCH_DOUBLES - 1 work
 * units or more are refused with KR_ERR_UNSUPPORTED.  All calls on one
 * device share its array: calls that overlap on one device (several
 * contexts or threads) race.  Blocks until the value is on the host. */
#define KR_NOALLOC_SCRATCH_DOUBLES 131072
int kr_sweep_solver_population_noalloc(kr_problem *prob, const int32_t *sdom_list, int32_t num, double *population);
/* Kernel::population (src/Kripke/Kernel/Population.cpp:74-101):
 * sum of w(d)*psi(d,g,z)*volume(z); deterministic fixed-order tree sum.
 * Blocks until the value is on the host. */
int kr_population(kr_problem *prob, double *out);

/* ---- transposed operators and parameter gradients ----------------------- */
/* No counterpart in the reference: the differentiable form of the source
 * iteration.  Each adjoint keeps its adjoint variable in the field of the
 * primal variable it belongs to (psibar in psi, rhsbar in rhs, ...), so an
 * adjoint problem registered with its own psi/rhs/phi/phi_out chunks (and
 * the forward problem's geometry, materials and quadrature) runs the chain
 * backwards with no copies:
 *   kr_ltimes_adjoint          psi      = L^T phi                (overwrite)
 *   kr_lplustimes_adjoint      phi_out  = (L+)^T rhs             (overwrite)
 *   kr_scattering_adjoint      phi     (+)= S^T phi_out
 *   kr_sweep_solver_adjoint    rhs      = H^-T psi  (all subdomains, the
 *                              transposed diamond-difference sweep: every
 *                              axis reversed, faces handed to the forward-
 *                              upwind subdomain, vacuum = zero adjoint inflow)
 * Per-zone isotropic group-diagonal scattering s(g,z), one array per PR
 * chunk in sigma_t's (sigt_zonal) layout, for the per-zone sigma_s parameter:
 *   kr_scattering_zonal        phi_out(0,g,z) (+)= s(g,z) phi(0,g,z), other moments (+)= 0
 *   kr_scattering_zonal_adjoint  its transpose (phi from phi_out)
 * Gradients into per-PR-chunk arrays in sigt_zonal layout (fwd and adj
 * share a context and a shape; fwd holds psi' = H^-1 rhs and phi, adj holds
 * rhsbar and phi_outbar of the same iteration):
 *   kr_grad_sigt        g(g,z) (+)= -sum over the chunk's subdomains, d of adj.rhs * fwd.psi
 *   kr_grad_sigs_zonal  g(g,z) (+)=  adj.phi_out(0,g,z) * fwd.phi(0,g,z)
 * Pointer arrays are host arrays of num_pr device pointers. */
int kr_ltimes_adjoint(kr_problem *prob);
int kr_lplustimes_adjoint(kr_problem *prob);
int kr_scattering_adjoint(kr_problem *prob, int32_t accumulate);
int kr_scattering_zonal(kr_problem *prob, const double *const *sigs_zonal, int32_t accumulate);
int kr_scattering_zonal_adjoint(kr_problem *prob, const double *const *sigs_zonal, int32_t accumulate);
int kr_sweep_solver_adjoint(kr_problem *prob);
/* No counterpart in the reference.  kr_sweep_solver_adjoint with adjoint
 * inflow on the domain boundary: where the forward flow leaves the domain
 * (kr_sdom_exits), the incoming adjoint faces are read from the subdomain's
 * own plane chunk ("i_plane"/"j_plane"/"k_plane"), which the caller fills
 * with the cotangent of the exiting faces.  After a forward
 * kr_sweep_solver, those chunks of the forward problem hold the exiting
 * faces themselves (SweepSolver.cpp:71-75: sweepSubdomain writes the
 * outgoing faces into the subdomain's planes and no downwind subdomain
 * takes them), so a loss on the exiting angular flux differentiates
 * through this call. */
int kr_sweep_solver_adjoint_seeded(kr_problem *prob);
/* exits[a] = 1 when subdomain sdom's downwind face along axis a (0 = i,
 * 1 = j, 2 = k) is the domain boundary (no downwind subdomain). */
int kr_sdom_exits(kr_problem *prob, int32_t sdom, int32_t *exits);
int kr_grad_sigt(kr_problem *fwd, kr_problem *adj, double *const *grad, int32_t accumulate);
int kr_grad_sigs_zonal(kr_problem *fwd, kr_problem *adj, double *const *grad, int32_t accumulate);

/* ---- differentiable source iteration: forward/backward pair -------------- */
/* No counterpart in the reference: K source iterations from psi = 0
 * (src/Kripke/SteadyStateSolver.cpp:44-104) as a function of per-zone
 * sigma_t(g,z) and isotropic group-diagonal scattering s(g,z) (one array per
 * PR chunk in sigt_zonal's layout), with reverse-mode gradients:
 *   phi_{n+1} = L H(sigma_t)^-1 L+ (S(s) phi_n + q),  phi_0 = 0,  output phi_K.
 * fwd and adj are two problems registered on one context from the same
 * inputs, each with its own psi, rhs, phi, phi_out and plane chunks (adj's
 * hold the adjoint variables).  Both must outlive the kr_diff.  num_fields
 * independent parameter sets share the two problems; each keeps its own
 * copy of sigma_t and s and phi checkpoints every checkpoint_every
 * iterations (memory per field: 2 parameter arrays + ceil(K/ck) phi
 * snapshots; plus ck-1 snapshots for the recomputation), and backward
 * recomputes the iterations between checkpoints.
 *
 * kr_diff_set_params   copies field f's parameters (num_pr device arrays
 *                      each); returns when they are copied.
 * kr_diff_forward      runs fields [first, first+num): phi_K of each into
 *                      phi_k[(f-first)*num_pr + c] and the exiting faces of
 *                      its last sweep into faces_out[(f-first)*num_exits + e]
 *                      (either array, or any entry, may be NULL: not
 *                      copied).  After the call the forward problem holds the
 *                      last field's phi_K and exiting faces.
 * kr_diff_backward     gradient of <phibar, phi_K> + <facebar, exiting faces>
 *                      w.r.t. sigma_t and s for fields [first, first+num),
 *                      overwriting grad_sigt/grad_s[(f-first)*num_pr + c];
 *                      phibar/facebar NULL (or a NULL entry) = zero
 *                      cotangent.  Needs the field's forward pass of its
 *                      current parameters (KR_ERR_INVALID otherwise); leaves
 *                      intermediate states in both problems' fields.
 * Exit e (kr_diff_exit) is subdomain sdom's plane chunk along axis (0 = i,
 * 1 = j, 2 = k) whose faces leave the domain, in (sdom, axis) order.
 * Both passes write each field's sigma_t into the two problems'
 * sigt_zonal chunks (the field the sweeps read).  Work is ordered on the
 * context stream; kr_synchronize before reading outputs elsewhere. */
typedef struct kr_diff kr_diff;
int kr_diff_create(kr_problem *fwd, kr_problem *adj, int32_t iterations, int32_t checkpoint_every,
                   int32_t num_fields, double source_strength, kr_diff **out);
int kr_diff_destroy(kr_diff *diff);
int kr_diff_num_exits(kr_diff *diff, int32_t *num_exits);
int kr_diff_exit(kr_diff *diff, int32_t exit, int32_t *sdom, int32_t *axis, int64_t *num_elem);
int kr_diff_set_params(kr_diff *diff, int32_t field, const double *const *sigt, const double *const *s);
int kr_diff_forward(kr_diff *diff, int32_t first, int32_t num, double *const *phi_k, double *const *faces_out);
int kr_diff_backward(kr_diff *diff, int32_t first, int32_t num, const double *const *phibar,
                     const double *const *facebar, double *const *grad_sigt, double *const *grad_s);

/* ---- autotuner --------------------------------------------------------- */
/* No counterpart in the reference.  Times the pipelined sweep (the dominant
 * kernel) over its shape candidates -- zones per lane of the row-scan kernel
 * or the systolic kernel's patch size, times three work-unit sizes -- on
 * this problem's own fields, installs the fastest, and caches the choice on
 * the context for every later problem of the same shape.  It runs the sweep,
 * so psi and the planes are overwritten: call it before the solve (or before
 * kr_batch_create).  The candidates are timed on the problem's current rhs:
 * after one iteration (nonzero sources and fluxes) the ranking is the one
 * the solve sees, while an all-zero rhs can rank them differently (the
 * device draws less power moving zeros).  choice (may be NULL) receives {kernel kind (as
 * kr_problem_sweep_kind reports it), zones per lane or patch columns, unit size in percent of the
 * default}. */
int kr_problem_autotune(kr_problem *prob, int32_t *choice);
/* No counterpart in the reference.  Installs one point of the autotuner's
 * search space for the pipelined sweep instead of timing them all:
 * v_max sets the row-scan kernel's zones per lane in the zone-fastest
 * nestings (0, 2 or 3 = two, the default; 1 = one; 4 = four where the rows,
 * of at most 64 zones, divide by four) or, in the other nestings, selects
 * the kernel (0 or 4 = systolic pencils, 2 = wave-uniform planes, 3 =
 * systolic pencils with the patches in k-fastest order and the faces
 * between k-patches in LDS, each where the shape fits it); a shape a choice
 * does not fit falls back to fewer zones per lane or to the plain systolic
 * kernel.  cap_div shrinks the systolic patch (both forms) or the
 * wave kernel's waves per workgroup (16 / cap_div);
 * r_scale_pct scales the pair groups per work unit.  kr_problem_sweep_kind
 * reports the kernel installed: 0 systolic, 1 row scan, 3 wave-uniform,
 * 4 systolic with the LDS k-face ring. */
int kr_problem_set_tuning(kr_problem *prob, int32_t v_max, int32_t cap_div, int32_t r_scale_pct);
int kr_problem_sweep_kind(kr_problem *prob, int32_t *kind);

/* ---- batched persistent launch ------------------------------------------ */
/* No counterpart in the reference: many independent problems (each a
 * registered kr_problem on one context) advanced together.  kr_batch_iterate
 * runs niter source iterations of every problem -- the on-device pipeline:
 * LTimes, scattering, source(strength), LPlusTimes, SweepSolver with the
 * population -- as ONE hipGraph launch of multi-problem kernels (one launch
 * covers the chunks or sweep units of every problem that takes the same
 * kernel variant), and returns the populations as
 * populations[problem*niter + iteration].  Each problem's results equal
 * those of its own iterations run one call at a time, bit for bit.  The
 * graph is captured on the first call and re-captured when niter changes, a
 * problem is re-tuned, or a problem's sigs changes between group-diagonal
 * and coupled; the problems' fields must not be re-registered while the
 * batch lives.  Member problems must outlive the batch: once one is
 * destroyed, kr_batch_iterate fails with KR_ERR_INVALID (kr_batch_destroy
 * still releases the batch).  Blocks until the values are on the host. */
typedef struct kr_batch kr_batch;
int kr_batch_create(kr_context *ctx, kr_problem *const *problems, const double *source_strength, int32_t n,
                    kr_batch **out);
int kr_batch_destroy(kr_batch *batch);
int kr_batch_iterate(kr_batch *batch, int32_t niter, double *populations);

/* ---- per-entry-point timing (hipEvents on the context stream) ---------- */
/* Names: LTimes, LPlusTimes, Scattering, Source, SweepSubdomain,
 * SweepSolver, Population, kConst.  kr_problem_timer synchronizes. */
int kr_problem_set_timing(kr_problem *prob, int32_t enable);
int kr_problem_timer(kr_problem *prob, const char *name, double *seconds, int64_t *count);
int kr_problem_timer_reset(kr_problem *prob);

/* ======================================================================= */
/* Host mirror: a from-scratch C++ DataStore/Set/Field + generateProblem +  */
/* SteadyStateSolver that drives the entry points above.                   */
/* ======================================================================= */

/* InputVariables (src/Kripke/InputVariables.h:19-46) plus the seeded
 * geometry of the benchmark configurations. */
typedef struct kr_input_vars {
  int32_t nx, ny, nz;
  int32_t num_directions;
  int32_t quad_num_polar, quad_num_azimuthal; /* 0,0 = dummy S2 set */
  int32_t num_groups;
  int32_t legendre_order;
  int32_t layout;                    /* kr_nesting */
  int32_t num_dirsets, num_groupsets;
  int32_t num_zonesets[3];
  int32_t num_material_subsamples;
  int32_t geometry;                  /* 0 = Kobayashi-3i (Space.cpp:142-177), 1 = seeded boxes in a 100 cm cube */
  uint64_t seed;
  double sigt[3], sigs[3];           /* used by geometry 0 */
  double source_strength;            /* Source.cpp:98 uses 1.0 */
  int32_t parallel_method;           /* InputVariables.h:42: 0 = sweep, 1 = Block Jacobi (--pmethod bj) */
} kr_input_vars;

typedef struct kr_host_problem kr_host_problem;

/* generateProblem (src/Kripke/Generate.cpp:24-94).  ctx == NULL builds the
 * host-side setup fields only (no device memory; solver calls fail). */
int kr_host_create(kr_context *ctx, const kr_input_vars *in, kr_host_problem **out);
int kr_host_destroy(kr_host_problem *hp);
int kr_host_num_sdoms(kr_host_problem *hp, int32_t *num_sdoms, int32_t *num_pr, int32_t *num_r);
kr_problem *kr_host_kernel_problem(kr_host_problem *hp);
/* Field access by reference name and chunk (element count returned in
 * *num_elem; dst may be NULL to query).  Integer fields convert to double. */
int kr_host_field_get(kr_host_problem *hp, const char *name, int32_t chunk, double *dst, int64_t *num_elem);
int kr_host_field_set(kr_host_problem *hp, const char *name, int32_t chunk, const double *src);
/* Device address and element count of a double field's chunk (device-
 * resident problems), for callers that move data on the device. */
int kr_host_field_device(kr_host_problem *hp, const char *name, int32_t chunk, void **ptr, int64_t *num_elem);
/* SteadyStateSolver (src/Kripke/SteadyStateSolver.cpp:24-115).
 * mode 0: the reference call sequence (kConst fills, accumulating entry
 *         points, host-side SweepComm queue over kr_sweep_subdomain).
 * mode 1: the on-device pipeline (fills folded into the kernels,
 *         kr_sweep_solver).  Both give the same phi and psi.
 * parallel_method 1 sweeps with Block Jacobi in both modes
 * (SteadyStateSolver.cpp:82-92 passes block_jacobi to SweepSolver).
 * kr_host_init_psi zeroes psi (SteadyStateSolver.cpp:39). */
int kr_host_init_psi(kr_host_problem *hp);
int kr_host_iterate(kr_host_problem *hp, int32_t mode, double *population);
int kr_host_solve(kr_host_problem *hp, int32_t niter, int32_t mode, double *populations);
/* No counterpart in the reference (its nesting is the --layout option,
 * src/Kripke/Generate.cpp:56-70 dispatches on it).  Picks the nesting for a
 * problem: builds it in each of the six nestings on ctx, tunes each one's
 * sweep shape (kr_problem_autotune), and times one on-device source
 * iteration (best of two after a warm-up); returns the fastest in *nesting
 * and its time in *ms (may be NULL).  Cached per context for problems of
 * the same extents, quadrature, groups, order and sets.  Nestings whose
 * shapes the kernels do not cover are skipped. */
int kr_host_autotune_nesting(kr_context *ctx, const kr_input_vars *in, int32_t *nesting, double *ms);

#ifdef __cplusplus
}
#endif
#endif /* KRIPKE_AMD_H */
