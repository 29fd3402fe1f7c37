/*
 * spimc.h -- C ABI of the MI355X action-evaluation and sampling engine.
 *
 * The engine replaces the hot path of simpimc's Action / Move / Observable
 * plug-ins for one species of Boltzmann particles in a cubic periodic box:
 * bead storage, the Kinetic action, the Bare and David pair actions with their
 * Ewald long-range (rho_k) part, the Bisect and DisplaceParticle Metropolis
 * moves, and the Energy, StructureFactor and PairCorrelation estimators.
 * Reference interfaces replaced are cited per entry point (paths relative to
 * the simpimc source tree).
 *
 * Conventions
 *   - Every entry point returns an int status: SPIMC_OK (0) or a negative
 *     SPIMC_ERR_* code.  No C++ exception crosses the boundary.  The text of the
 *     last error of the calling thread is available from spimc_last_error().
 *   - All buffers are caller-owned host memory; device memory is owned by the
 *     context.  A context owns one HIP stream; calls on one context are not
 *     re-entrant.  Calls return after the device work they issue has finished.
 *   - A context holds one or more independent replicas (simpimc's MPI clones,
 *     framework/framework_class.h:44-53).  Replicas may differ in every size.
 *   - Bead positions cross the boundary as slice-major structure-of-arrays:
 *     r[d][b][p], d in 0..2, b in 0..n_bead-1, p in 0..n_part-1 (FP64).
 *   - rho_k crosses as rho[b][k][2] (real, imaginary), FP64.
 *   - Everything is FP64.
 */
#ifndef SPIMC_H_
#define SPIMC_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SPIMC_ABI_VERSION 6

/* status codes */
#define SPIMC_OK 0
#define SPIMC_ERR_ARG (-1)         /* invalid argument */
#define SPIMC_ERR_UNSUPPORTED (-2) /* valid reference input this engine does not cover */
#define SPIMC_ERR_HIP (-3)         /* HIP runtime error */
#define SPIMC_ERR_STATE (-4)       /* call out of sequence (e.g. no proposal set) */
#define SPIMC_ERR_NOMEM (-5)       /* device allocation failed */
#define SPIMC_ERR_NODEVICE (-6)    /* no HIP device */

/* pair action families (actions/actions.h:13-35 "type" strings) */
#define SPIMC_PAIR_NONE 0
#define SPIMC_PAIR_BARE 1  /* "BarePairAction"  pair_action/bare_pair_action_class.h */
#define SPIMC_PAIR_DAVID 2 /* "DavidPairAction" pair_action/david_pair_action_class.h */
#define SPIMC_PAIR_ILKKA 3 /* "IlkkaPairAction" pair_action/ilkka_pair_action_class.h */

/* action selectors */
#define SPIMC_ACTION_KINETIC 0
#define SPIMC_ACTION_PAIR 1

/* path modes (data_structures/bead_class.h:7-8) */
#define SPIMC_MODE_OLD 0
#define SPIMC_MODE_NEW 1

/* gradient terms (bit mask) */
#define SPIMC_GRAD_KINETIC 1u
#define SPIMC_GRAD_PAIR_SHORT 2u
#define SPIMC_GRAD_PAIR_LONG 4u
#define SPIMC_GRAD_ALL 7u

/* energy estimators (events/observables/energy_class.h:22-116) */
#define SPIMC_ENERGY_THERMAL 0
#define SPIMC_ENERGY_VIRIAL_COMPAT 1   /* reference virial: FD pair gradient, no Ewald term */
#define SPIMC_ENERGY_VIRIAL_ADJOINT 2  /* virial from the analytic adjoint, Ewald included */

/* observables (bit mask) */
#define SPIMC_OBS_ENERGY 1u
#define SPIMC_OBS_SK 2u
#define SPIMC_OBS_GR 4u

/* sweep flags */
#define SPIMC_SWEEP_GRAPH 1u   /* replay a captured hipGraph of the sweep */
#define SPIMC_SWEEP_ADJOINT 2u /* compute dS/dr of every bead after the moves */
#define SPIMC_SWEEP_ENERGY 4u  /* accumulate the configured energy estimator */
#define SPIMC_SWEEP_SK 8u      /* accumulate S(k) */
#define SPIMC_SWEEP_GR 16u     /* accumulate g(r) */

/* kernels timed by spimc_kernel_timing */
#define SPIMC_KERNEL_PROPOSE 0
#define SPIMC_KERNEL_PAIR 1
#define SPIMC_KERNEL_RHO 2
#define SPIMC_KERNEL_DECIDE 3
#define SPIMC_KERNEL_ADJOINT 4
#define SPIMC_KERNEL_ESTIMATORS 5
#define SPIMC_KERNEL_COUNT 6

/* timing modes of spimc_timing_enable (may be or-ed) */
#define SPIMC_TIMING_CLOCK 1  /* device wall-clock spans of the pair and rho_k kernels; no events */
#define SPIMC_TIMING_EVENTS 2 /* start/stop events written by every kernel's dispatch */

/*
 * One replica's system.  Field groups follow the reference XML elements:
 *   <System n_d n_bead beta L>                 (data_structures/path_class.h:25-69)
 *   <Species n_part lambda>                    (data_structures/species_class.h:48-98)
 *   <Action type="Kinetic" n_images>           (actions/single_action/kinetic_class.h:28-32)
 *   <Action type="BarePairAction"|"DavidPairAction"|"IlkkaPairAction" n_order max_level use_long_range>
 *                                              (actions/pair_action/pair_action_class.h:205-238)
 *   <Move type="Bisect" n_level n_images>      (events/moves/single_species_move/bisect/bisect_class.h:167-190)
 *   <Move type="DisplaceParticle" step_size>   (events/moves/single_species_move/displace_particle_class.h:93)
 * Pair-action tables are passed as the knot values the reference loaders hand
 * to einspline (after their own assembly), so that the loaders stay host code.
 */
typedef struct spimc_system_desc {
    /* <System> */
    uint32_t n_d;    /* must be 3 */
    uint32_t n_bead; /* M */
    double beta;     /* tau = beta / M */
    double L;        /* cubic box side; periodic */
    /* <Species> */
    uint32_t n_part; /* N */
    double lambda;   /* hbar^2 / 2m */
    /* k-vectors: integer triples n, stored vector k = (2 pi / L) n.  The set is
     * the one KSpace::Setup keeps (spimc_kspace_build); order is the caller's. */
    uint32_t n_k;
    const int32_t* k_n; /* [n_k][3] */
    /* <Action type="Kinetic"> */
    int32_t kinetic_n_images;
    /* <Move type="Bisect"> */
    uint32_t bisect_n_level;
    int32_t bisect_n_images;
    /* <Move type="DisplaceParticle"> */
    double displace_step;
    /* pair action */
    int32_t pair_type; /* SPIMC_PAIR_* */
    int32_t pair_use_long_range;
    uint32_t pair_max_level; /* must be 0 */
    /* BarePairAction knots (v/diag/{r,v_r,r_long,v_long_r}) */
    int32_t bare_is_coulomb;
    uint32_t bare_n_r;
    const double* bare_r;
    const double* bare_v_r;
    uint32_t bare_n_r_long;
    const double* bare_r_long;
    const double* bare_v_long_r;
    /* DavidPairAction knots: n_order, grid, and the (n_val+1) x n_grid arrays
     * u_kj and du_kj_dbeta for tau, component 0 = potential (component-major). */
    uint32_t david_n_order;
    uint32_t david_n_grid;
    const double* david_grid;
    const double* david_u_kj;
    const double* david_du_kj;
    /* long range: per stored k-vector v_long_k (Bare: as read; David: u_k/volume),
     * and the raw k=0 and r=0 constants before the homologous scaling of
     * bare_pair_action_class.h:89-95 / david_pair_action_class.h:340-361. */
    const double* pair_v_long_k; /* [n_k] */
    double pair_v_long_k_0;
    double pair_v_long_r_0;
    /* IlkkaPairAction knots (ilkka_pair_action_class.h:253-432), every spline a
     * natural cubic one on its own general grid:
     *   u/off_diag: U(x, y) on x_u[n_x_u] x y_u[n_y_u], values u_xy[n_y_u][n_x_u]
     *     (x fastest, the memory order of the reference's mat(n_x, n_y));
     *   du/off_diag: dU/dbeta(x, y) likewise;
     *   v/diag: v_r on r_v (the potential; ilkka_v_long_r on ilkka_r_v_long is its
     *     long-range part);
     *   u/diag, du/diag: u_long_r, du_long_r on their own grids (use_long_range).
     * Per stored k-vector: ilkka_u_long_k (action), ilkka_du_long_k (dU/dbeta);
     * pair_v_long_k holds v/diag v_long_k (potential).  The raw k=0 and r=0
     * constants of du/diag are ilkka_du_long_k_0, ilkka_du_long_r_0; those of
     * v/diag are pair_v_long_k_0, pair_v_long_r_0. */
    uint32_t ilkka_n_x_u, ilkka_n_y_u;
    const double* ilkka_x_u;
    const double* ilkka_y_u;
    const double* ilkka_u_xy;
    uint32_t ilkka_n_x_du, ilkka_n_y_du;
    const double* ilkka_x_du;
    const double* ilkka_y_du;
    const double* ilkka_du_xy;
    uint32_t ilkka_n_r_v;
    const double* ilkka_r_v;
    const double* ilkka_v_r;
    uint32_t ilkka_n_r_v_long;
    const double* ilkka_r_v_long;
    const double* ilkka_v_long_r;
    uint32_t ilkka_n_r_u_long;
    const double* ilkka_r_u_long;
    const double* ilkka_u_long_r;
    uint32_t ilkka_n_r_du_long;
    const double* ilkka_r_du_long;
    const double* ilkka_du_long_r;
    const double* ilkka_u_long_k;  /* [n_k] */
    const double* ilkka_du_long_k; /* [n_k] */
    double ilkka_du_long_k_0;
    double ilkka_du_long_r_0;
} spimc_system_desc;

typedef struct spimc_ctx spimc_ctx;

/* ---- library ---------------------------------------------------------- */

/* ABI version and the text of the calling thread's last error. */
int spimc_abi_version(void);
const char* spimc_last_error(void);

/* k-vector set of KSpace::Setup / KSpace::Include (data_structures/k_space_class.h:17-80):
 * integer triples n != 0 with |2 pi n / L|^2 < k_cut^2 in the half space
 * (n0 > 0) or (n0 == 0 and n1 > 0) or (n0 == n1 == 0 and n2 > 0).  Writes up
 * to cap triples (lexicographic in n0, n1, n2) and returns the count in *n_k.
 * Host only; needs no device. */
int spimc_kspace_build(double L, double k_cut, int32_t* k_n, uint32_t cap, uint32_t* n_k);

/* ---- context ---------------------------------------------------------- */

/* Creates a context with n_replicas replicas on the current HIP device.
 * Replaces Simulation::Simulation (framework/simulation_class.h:16-39) for the
 * hot path: allocates slice-major SoA positions, per-slice rho_k (two buffers),
 * splines and work lists in HBM.  seed keys the device RNG of every replica. */
int spimc_ctx_create(const spimc_system_desc* descs, uint32_t n_replicas, uint64_t seed, spimc_ctx** out);
int spimc_ctx_destroy(spimc_ctx* ctx);
int spimc_ctx_num_replicas(const spimc_ctx* ctx, uint32_t* out);

/* Path storage: Bead::SetR/StoreR (data_structures/bead_class.h:58-104) for the
 * whole path.  Upload also rebuilds rho_k of every slice (Species::InitRhoK,
 * data_structures/species_class.h:380-403). */
int spimc_path_upload(spimc_ctx* ctx, uint32_t replica, const double* r);
int spimc_path_download(spimc_ctx* ctx, uint32_t replica, double* r);
/* Committed per-slice rho_k (Species::GetRhoK in OLD mode, species_class.h:428). */
int spimc_rho_k_download(spimc_ctx* ctx, uint32_t replica, double* rho);
/* Rebuild rho_k of every slice from the positions (Species::InitRhoK). */
int spimc_rho_k_rebuild(spimc_ctx* ctx, uint32_t replica);

/* ---- Action interface (actions/action_class.h:37-64) ------------------ */

/* NEW-mode positions for particle `particle`, slices b_first .. b_first+n-1
 * (mod n_bead), r_new[n][3].  Mirrors Bead::SetR in NEW mode. */
int spimc_set_proposal(spimc_ctx* ctx, uint32_t replica, uint32_t particle, uint32_t b_first, uint32_t n_beads, const double* r_new);
/* Move::Accept / Move::Reject for an interface proposal (bisect_class.h:24-36,127-139):
 * accept stores NEW positions and NEW rho_k of the touched slices, reject drops them. */
int spimc_accept_proposal(spimc_ctx* ctx, uint32_t replica);
int spimc_reject_proposal(spimc_ctx* ctx, uint32_t replica);

/* Action::GetAction(b0, b1, particles, level) (action_class.h:43) of the
 * Kinetic (kinetic_class.h:105-122) or pair action (pair_action_class.h:267-302)
 * in OLD or NEW mode.  NEW mode on the pair action with long range also runs
 * Species::UpdateRhoK (species_class.h:406-425) for the proposal's slices.
 * particles holds at most one index (Boltzmann moves move one particle). */
int spimc_get_action(spimc_ctx* ctx, uint32_t replica, int action, int mode, uint32_t b0, uint32_t b1,
                     const uint32_t* particles, uint32_t n_particles, uint32_t level, double* out);
/* Action::GetActionLaplacian(b0, b1, particles, level) (action_class.h:49) in
 * OLD or NEW mode, arguments as spimc_get_action.  Kinetic
 * (kinetic_class.h:147-165): n_d 4 / (4 lambda tau 2^level) per step of the
 * particle's walk.  Pair action (pair_action_class.h:336-363): for the window
 * beads b, the central second difference (eps 1e-4, CalcLaplacianU :167-199)
 * of the links (b, b + 2^level) and (b, b - 2^level) with every other
 * particle in each coordinate of bead (p, b); 0 above level 0 (max_level); the
 * long-range part is left out, as in the reference.  Its only reference caller
 * is the ContactDensity observable. */
int spimc_get_action_laplacian(spimc_ctx* ctx, uint32_t replica, int action, int mode, uint32_t b0, uint32_t b1,
                               const uint32_t* particles, uint32_t n_particles, uint32_t level, double* out);

/* spimc_get_action without allocation.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, results and errors as spimc_get_action, whose kernels it shares;
 * its scratch (long-range partials, the result, work items) is a fixed device
 * array of SPIMC_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends,
 * instead of context buffers grown on demand.  It calls no hipMalloc or
 * hipFree.  Limits: n_bead + 8 <= SPIMC_NOALLOC_SCRATCH_BYTES / 32 and
 * n_bead * (k chunks) <= 3 * SPIMC_NOALLOC_SCRATCH_BYTES / 32; larger replicas
 * get SPIMC_ERR_ARG.  Refused (SPIMC_ERR_STATE) while SPIMC_TIMING_EVENTS is
 * on.  All contexts on one GPU share the array: calls that overlap on one GPU
 * (from several threads or contexts) race. */
#define SPIMC_NOALLOC_SCRATCH_BYTES (256u * 1024u)
int spimc_get_action_noalloc(spimc_ctx* ctx, uint32_t replica, int action, int mode, uint32_t b0, uint32_t b1,
                             const uint32_t* particles, uint32_t n_particles, uint32_t level, double* out);

/* Action::DActionDBeta (action_class.h:37): kinetic_class.h:35-45,
 * pair_action_class.h:241-264, committed path. */
int spimc_d_action_d_beta(spimc_ctx* ctx, uint32_t replica, int action, double* out);
/* Action::Potential (action_class.h:52): pair_action_class.h:369-395. */
int spimc_potential(spimc_ctx* ctx, uint32_t replica, int action, double* out);
/* Kinetic::VirialEnergy(window) (kinetic_class.h:48-102). */
int spimc_kinetic_virial_energy(spimc_ctx* ctx, uint32_t replica, uint32_t window, double* out);

/* Per-bead action gradient out[b][p][3] of the committed path.
 *   compat = 1: the reference's PairAction::GetActionGradient(b, b+1, {p}, 0)
 *               (pair_action_class.h:305-337) for every bead.  SPIMC_GRAD_PAIR_SHORT:
 *               its CalcGradientU on the links (b, b+1) and (b, b-1) -- a central
 *               difference (eps 1e-4) of the link action for Bare and David, the
 *               analytic form of ilkka_pair_action_class.h:172-222 for Ilkka.
 *               SPIMC_GRAD_PAIR_LONG: CalcGradientULong once per particle pair --
 *               zero for Bare and David, ilkka_pair_action_class.h:231-249 for Ilkka.
 *               `terms` must not hold SPIMC_GRAD_KINETIC.
 *   compat = 0: analytic adjoint dS/dr of the terms selected in `terms`
 *               (kinetic, short-range pair, Ewald long range) in one pass.
 *               Replaces the numeric gradients feeding energy_class.h:79-94. */
int spimc_action_gradient(spimc_ctx* ctx, uint32_t replica, uint32_t terms, int compat, double* out);

/* ---- Move interface (events/moves/move_class.h:21-77) ----------------- */

/* Bisect::Reset with adaptive="1" (events/moves/single_species_move/bisect/
 * bisect_class.h:143-151): sets one replica's bisection level for the stages
 * and sweeps that follow (1 <= n_level <= min(6, floor(log2 n_bead))).
 * Synchronises the device. */
int spimc_set_bisect_level(spimc_ctx* ctx, uint32_t replica, uint32_t n_level);

/* Replay draws for one replica, in reference call order (post-transform):
 *   Bisect::Attempt (bisect_class.h:39-124): p, b0 (integers as doubles), then per
 *     level from n_level-1 down to 0: 3 standard normals per midpoint in path
 *     order, then one uniform in (0,1]; the stream stops after the uniform of a
 *     rejecting level.
 *   DisplaceParticle::Attempt (displace_particle_class.h:28-72): p, dx, dy, dz, uniform.
 * draws == NULL switches the replica back to its device RNG. */
int spimc_tape_set(spimc_ctx* ctx, uint32_t replica, const double* draws, uint64_t n_draws);
int spimc_tape_cursor(spimc_ctx* ctx, uint32_t replica, uint64_t* cursor);

/* n_attempts[r] consecutive Move::DoEvent calls of Bisect (resp. DisplaceParticle)
 * on every replica r (n_attempts == NULL: n_uniform each).  accepted (optional)
 * receives one byte per attempt, replica-major with row length max_r n_attempts[r]. */
int spimc_bisect_stage(spimc_ctx* ctx, const uint32_t* n_attempts, uint32_t n_uniform, uint8_t* accepted);
int spimc_displace_stage(spimc_ctx* ctx, const uint32_t* n_attempts, uint32_t n_uniform, uint8_t* accepted);
/* Move counters n_attempt / n_accept (move_class.h:61-77) since creation:
 * out[0..3] = bisect attempts, bisect accepts, displace attempts, displace accepts. */
int spimc_move_counters(spimc_ctx* ctx, uint32_t replica, uint64_t* out);

/* Action trace of the move stages: with it on, every attempt of a Bisect,
 * DisplaceParticle or PermBisectIterative stage records the terms its level-0
 * Metropolis decision adds up -- the values the reference's Attempt takes from
 * Action::GetAction in OLD and NEW mode at level 0 (bisect_class.h:100-114,
 * displace_particle_class.h:52-68, perm_bisect_iterative_class.h:194-213):
 *   out[i][0..7] = kinetic OLD, kinetic NEW, short-range pair OLD, NEW,
 *                  long-range (Ewald k-space) pair OLD, NEW,
 *                  log sample ratio at level 0, action change of level 1
 * for attempt i of the replica's last stage (NaN for an attempt that made no
 * level-0 decision: rejected at a coarser level, or its cycle walk stopped).
 * Short-range pair terms sum PairAction::CalcU over the window links and
 * partners (pair_action_class.h:279-290); long-range terms are CalcULong
 * (bare_pair_action_class.h:153-172, david_pair_action_class.h:105-123) on the
 * updated rho_k.  Enabling synchronises the device.  Read n_attempts <= the
 * replica's attempts in the last stage.  Diagnostic; no counterpart in the
 * reference, which keeps these values in locals. */
int spimc_action_trace_enable(spimc_ctx* ctx, int enable);
int spimc_action_trace_read(spimc_ctx* ctx, uint32_t replica, uint32_t n_attempts, double* out);

/* ---- Permutations (events/moves/single_species_move/bisect/perm_bisect/) --
 * A replica's path topology is the permutation sigma at the wrap: the bead after
 * (p, n_bead - 1) is (sigma[p], 0), where PermBisect::AssignParticleLabels
 * (perm_bisect_class.h:47-56) leaves every permutation.  The kinetic action,
 * the bisection walks, Kinetic::DActionDBeta / VirialEnergy and the adjoint's
 * kinetic term follow it; the pair action and rho_k stay per label and slice
 * (Path::DrDrpDrrp, path_class.h:137-150).  Identity at creation. */

/* PermBisectIterative's attributes n_level and epsilon (perm_bisect_class.h:104,
 * bisect_class.h:168) for one replica (defaults: the descriptor's bisect level,
 * 1e-100).  Synchronises the device. */
int spimc_perm_configure(spimc_ctx* ctx, uint32_t replica, uint32_t n_level, double epsilon);
/* n_attempts[r] consecutive Move::DoEvent calls of PermBisectIterative
 * (perm_bisect_iterative_class.h:113-223, perm_bisect_class.h:59-90) on every
 * replica; arguments as spimc_bisect_stage.  Replay draws per attempt: bead0, the
 * first particle (integers as doubles); per walk step the continue uniform and,
 * unless the walk stops there, the select uniform; then per level the chains'
 * normals chain by chain (3 per midpoint in path order) and one uniform.
 * Replicas hold at most 4096 particles. */
int spimc_perm_bisect_stage(spimc_ctx* ctx, const uint32_t* n_attempts, uint32_t n_uniform, uint8_t* accepted);
/* PermBisectTable's attribute n_perm_part (perm_bisect_table_class.h:235) for
 * one replica, 1 <= n_perm_part <= min(n_part, 6), at most 1024 particles and
 * 2^28 cycle-list entries (n_part! / (n_part - n_perm_part)!).  balance selects
 * the final table check (perm_bisect_table_class.h:216-225): 0 the reference's,
 * accept if perm_tot_0 / perm_tot_1 >= u; 1 perm_tot_0 / (perm_tot_1 w) >= u,
 * w the selected entry's weight, which detailed balance asks for when
 * n_perm_part < n_part.  Uses spimc_perm_configure's n_level and epsilon.
 * Synchronises the device. */
int spimc_perm_table_configure(spimc_ctx* ctx, uint32_t replica, uint32_t n_perm_part, uint32_t balance);
/* n_attempts[r] consecutive Move::DoEvent calls of PermBisectTable
 * (perm_bisect_table_class.h:122-228, perm_bisect_class.h:59-90) on every
 * replica; arguments as spimc_bisect_stage.  The cycle list is every
 * permutation of n_perm_part particles (lexicographic) crossed with every
 * n_perm_part-subset of the particles (lexicographic), permutation-major.
 * Replay draws per attempt: bead0 (an integer as a double), the selection
 * uniform u (the entry whose running total of weights first exceeds u times
 * their total), per level the chains' normals chain by chain (3 per midpoint
 * in path order) and one uniform, then, after an accepted level 0 and when
 * n_perm_part < n_part, the table check's uniform.  Counters as
 * spimc_perm_counters, at cycle length n_perm_part. */
int spimc_perm_table_stage(spimc_ctx* ctx, const uint32_t* n_attempts, uint32_t n_uniform, uint8_t* accepted);
/* sigma[p] for p < n_part: PathDump's permutation row 1 (path_dump_class.h:43-47);
 * upload is the Restart path of species_class.h:366-374.  Upload checks that
 * sigma is a permutation. */
int spimc_perm_download(spimc_ctx* ctx, uint32_t replica, uint32_t* sigma);
int spimc_perm_upload(spimc_ctx* ctx, uint32_t replica, const uint32_t* sigma);
/* PermBisectIterative counters since creation: out[0] attempts, out[1] accepts
 * (Move::DoEvent), out[2 + c] perm_attempt and out[2 + n_part + c] perm_accept of
 * cycle length c + 1 (perm_bisect_class.h:25-26,66-87).  out holds 2 + 2 n_part. */
int spimc_perm_counters(spimc_ctx* ctx, uint32_t replica, uint64_t* out);

/* ---- Observable interface (events/observables/observable_class.h:72-104) */

/* Energy: estimator SPIMC_ENERGY_*, virial window (0 = n_bead).
 * StructureFactor: k_cut (structure_factor_class.h:15-32).
 * PairCorrelation: r_min, r_max, n_r (pair_correlation_class.h:15-28, LinearGrid
 * observable_class.h:20-63); r_max <= 0 selects L/2. */
int spimc_obs_configure(spimc_ctx* ctx, int energy_estimator, uint32_t virial_window, double sk_k_cut,
                        double gr_r_min, double gr_r_max, uint32_t gr_n_r);
/* Observable::Accumulate of the selected observables on every replica. */
int spimc_obs_accumulate(spimc_ctx* ctx, uint32_t which);
/* Raw accumulators of one replica since the last reset (what Write() normalises):
 *   SPIMC_OBS_ENERGY: out[0] = kinetic, out[1] = pair (sums of per-action energies)
 *   SPIMC_OBS_SK:     out[n_k]  sum over measurements and slices of |rho_k|^2 (k below k_cut)
 *   SPIMC_OBS_GR:     out[n_r]  histogram counts
 * *n_measure receives the number of accumulations. */
int spimc_obs_read(spimc_ctx* ctx, uint32_t replica, uint32_t which, double* out, uint64_t* n_measure);
int spimc_obs_reset(spimc_ctx* ctx, uint32_t which);

/* ---- batched sweep ---------------------------------------------------- */

/* n_sweeps sweeps of every replica: N*M/2^n_level bisect attempts, then N
 * displace attempts (of that replica's N, M), then per `flags` the adjoint
 * gradient and the estimators.  With SPIMC_SWEEP_GRAPH the rounds are replayed
 * from hipGraphs captured on first use. */
int spimc_sweep(spimc_ctx* ctx, uint32_t n_sweeps, uint32_t flags);
/* Bead updates performed so far (Bisect: 2^l - 1 per attempt, Displace: M). */
int spimc_bead_updates(spimc_ctx* ctx, uint64_t* out);
/* Launch-shape autotuning.  For each candidate shape (rho_k kernel beside the
 * pair kernel or after it, rho_k workgroups per CU, pair workgroup size; and,
 * when every replica has N M <= 16384 on unpermuted paths, the persistent
 * stage kernel that runs a replica's whole stage in one workgroup, with 256,
 * 512 or 1024 threads; and, for at most 256 replicas, rounds of one launch
 * each, for bisection or for both moves, whose last workgroup per replica
 * decides the replica's proposal and makes its next one) it runs one
 * bisection and one displace attempt per
 * replica untimed, then times n_rounds bisection attempts followed by
 * n_displace displace attempts (the sweep's ratio, N M / 2^l : N, for replica
 * 0); it keeps the fastest shape for later stages and drops captured graphs.
 * The stages are ordinary moves of the chains (call it during equilibration);
 * the shape never changes results beyond rounding.  chosen (optional, 7 ints):
 * fork, rho_k workgroups per CU, pair workgroup size, number of shapes tried,
 * n_displace, persistent stage workgroup size (0: rounds of launches),
 * one-launch rounds for bisection (1), for both moves (2), or none (0).  No
 * counterpart in the reference. */
int spimc_autotune(spimc_ctx* ctx, uint32_t n_rounds, int32_t* chosen);

/* ---- measurement ------------------------------------------------------ */

/* Timing of the move kernels (0 disables).  SPIMC_TIMING_CLOCK: each pair and
 * rho_k launch records its first block start and last block end on the GPU's
 * constant-rate wall clock and the round's decide kernel sums the spans, so the
 * timed run is the same program as an untimed one (also under graph replay).
 * SPIMC_TIMING_EVENTS: every launch carries dispatch start/stop events
 * (hipExtLaunchKernel); these flush caches around each kernel and slow the run.
 * No counterpart in the reference (measurement only). */
int spimc_timing_enable(spimc_ctx* ctx, int mode);
/* Total milliseconds, launches and algorithmic bytes (rho_k: 32 B per
 * (slice, k) update; pair: link evaluations) of one kernel class since the last
 * reset.  With SPIMC_TIMING_CLOCK the pair and rho_k classes report the device
 * clock spans of every launch (bytes count only the work done). */
int spimc_kernel_timing(spimc_ctx* ctx, int kernel, double* total_ms, uint64_t* launches, double* bytes);
int spimc_timing_reset(spimc_ctx* ctx);

#ifdef __cplusplus
}
#endif

#endif /* SPIMC_H_ */
