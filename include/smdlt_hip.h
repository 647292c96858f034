/*
 * smdlt_hip.h -- C ABI of the MI355X (gfx950) timestep hot path for bead-spring
 * Brownian dynamics with Ewald-summed RPY hydrodynamics, flowing-box remaps and
 * charge-hopping Monte Carlo.
 *
 * The reference has no plugin or FFI layer: its hot path is a set of
 * `void f(void)` functions sharing globals (S/main.h:5-65, S/Parameters.h:40-201;
 * S/ = shear/src/, X/ = extensional/src/).  Each entry point below names the
 * reference function whose body it replaces, so a maintainer can swap a body for
 * a call.  Plain C types only; host arrays belong to the caller and are copied.
 *
 * Conventions
 *  - Return 0 on success, a negative SMDLT_E_* code otherwise (the reference
 *    prints and exit(1)s instead).  smdlt_strerror() names the code.
 *  - A context owns its device memory and one HIP stream; calls on one context
 *    are serialised on that stream and are not re-entrant.
 *  - Bead arrays are concatenated over replicas in replica order; replica r has
 *    nb*nc beads, chain-major like the reference (bead = chain*nb + j).
 *  - Positions are SoA doubles (the reference's rx, ry, rz); vectors (forces,
 *    noise, velocities) are AoS 3N doubles (the reference's R[3N] layout).
 */
#ifndef SMDLT_HIP_H
#define SMDLT_HIP_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define SMDLT_OK 0
#define SMDLT_E_ARG (-1)             /* bad argument or size */
#define SMDLT_E_HIP (-2)             /* a HIP runtime call failed */
#define SMDLT_E_FENE_OVERSTRETCH (-3)/* S/main26.c:553-558 */
#define SMDLT_E_LANCZOS_NOCONV (-4)  /* tolerance not reached within max iterations */
#define SMDLT_E_NLIST_OVERFLOW (-5)  /* neighbour list capacity exceeded (retried internally) */
#define SMDLT_E_GEOMETRY (-6)        /* cutoff too large for the (deformed) box */
#define SMDLT_E_STATE (-7)           /* call order violated (e.g. no positions uploaded) */
#define SMDLT_E_NOMEM (-8)
#define SMDLT_E_MC_STREAM (-9)       /* MC ran out of host-fed uniforms */
#define SMDLT_E_NONFINITE (-10)      /* a step produced a non-finite displacement; it was not applied (positions
                                        stay finite) and the step's other beads moved as usual */

/* Flow types: the reference's -o flag and its two programs. */
#define SMDLT_FLOW_EQUILIBRIUM 0
#define SMDLT_FLOW_SHEAR 1       /* Lees-Edwards deforming brick (S/) */
#define SMDLT_FLOW_EXTENSION 2   /* planar extension, Kraynik-Reinelt lattice (X/) */

/* Mobility precision. */
#define SMDLT_PREC_MIXED 0  /* FP32 pair/grid/FFT sums, compensated per-bead sums, FP64 Lanczos */
#define SMDLT_PREC_FP64 1   /* every kernel in FP64 */

/* Brownian term. */
#define SMDLT_NOISE_LANCZOS 0 /* M^(1/2) z by Krylov (Lanczos) */
#define SMDLT_NOISE_TEA 1     /* the reference's CA-TEA product (S/main26.c:787-804); see smdlt_set_tea */
#define SMDLT_NOISE_FREE_DRAINING 2 /* no hydrodynamics: updateFD (S/main26.c:731-744), M = I */

/* MC modes. */
#define SMDLT_MC_REFERENCE 0 /* chargeHop's state machine and acceptance (S/main26.c:1926-2390) */
#define SMDLT_MC_MH_UNIFORM 1/* exact Metropolis-Hastings, uniform hop proposal within hop radius */
#define SMDLT_MC_MH_LEARNED 2/* exact Metropolis-Hastings, learned (MLP) proposal */

/* Per-replica parameters: the reference's CLI flags and Input.txt keys
 * (S/Initialization.c:52-174) plus the mobility settings of this build.  A zero
 * in an "auto" field lets smdlt_create pick the value (autotuner). */
typedef struct smdlt_params {
  /* chain set and box */
  int nb, nc;              /* -b beads per chain, -a chains */
  double L;                /* cubic box edge (initBox, S/Initialization.c:176-215) */
  int flow_type;           /* -o */
  double edot;             /* -f strain rate */
  double dt;               /* -d */
  /* Input.txt */
  int spring_fene;         /* spring = FENE (1) or Hookean (0) */
  double epsilon, kappab, kappas, qmax;
  double lambda_b;         /* Bjerrum length */
  /* charges */
  double lambda_d;         /* -l Debye length */
  double barrier, barrier2;/* -x, -y */
  int lj_ext_form;         /* 0: shear LJ 12eps(r6^2-r6), rc=5; 1: extensional WCA 48eps form */
  int pbc_compat;          /* 1: reproduce the shear +-delx wrap shift (S/main26.c:1412-1423) for parity
                              runs; 0 (default): the exact Lees-Edwards wrap */
  /* mobility (this build) */
  int precision;           /* SMDLT_PREC_* */
  int noise_mode;          /* SMDLT_NOISE_* */
  double ewald_alpha;      /* splitting parameter xi [1/a]; 0 = auto */
  double rc_real;          /* real-space cutoff [a]; 0 = auto */
  double skin;             /* Verlet skin of the real-space list [a]; 0 = 1.0 */
  int grid[3];             /* PME grid along a1,a2,a3 (multiples of 8 of the form 2^a 3^b 5^c, 16..512); 0 = auto */
  int spline_order;        /* B-spline order 4..8; 0 = auto */
  double mobility_tol;     /* target relative M.F error for the auto choices; 0 = 1e-5 */
  double lanczos_tol;      /* 0 = 1e-6 */
  int lanczos_maxit;       /* 0 = 30 */
  uint64_t seed;           /* device noise stream seed for this replica */
  /* Metropolis-Hastings hop modes (SMDLT_MC_MH_*) */
  double hop_radius;       /* candidate beads within this distance [a]; 0 = 3.0 */
  double feature_radius;   /* neighbourhood of the proposal features [a]; 0 = 4.0 */
  double proposal_floor;   /* learned mode: weight of the uniform part of q; 0 = 0.05 */
} smdlt_params;

typedef struct smdlt_ctx smdlt_ctx;

/* Library version string. */
const char* smdlt_version(void);
/* Name of an error code. */
const char* smdlt_strerror(int code);
/* Fill p with the reference defaults (S/Initialization.c:30-48, S/Input.txt). */
void smdlt_default_params(smdlt_params* p);

/* Create a context for n_replicas independent boxes (allocate + initLattice +
 * initVerlet: S/Initialization.c:217-457).  device = HIP device ordinal. */
int smdlt_create(const smdlt_params* params, int n_replicas, int device, smdlt_ctx** out);
void smdlt_destroy(smdlt_ctx* ctx);
int smdlt_num_beads(const smdlt_ctx* ctx, int replica); /* replica = -1: total */

/* Effective mobility settings chosen for a replica: alpha, rc, skin, K1, K2, K3,
 * order, the number of real-space bead-pair list entries after the last build
 * and (mixed precision) the number of cluster-pair entries. */
int smdlt_get_mobility_config(const smdlt_ctx* ctx, int replica, double out[9]);

/* Host <-> device state.  Positions SoA over all beads (reference rx, ry, rz). */
int smdlt_upload_positions(smdlt_ctx* ctx, const double* rx, const double* ry, const double* rz);
int smdlt_download_positions(smdlt_ctx* ctx, double* rx, double* ry, double* rz);
/* Charges: per replica nq[r] charges; bead index of each (Charge_track, local to
 * the replica) and its sign (+1 in the reference; -1 allowed). */
int smdlt_upload_charges(smdlt_ctx* ctx, const int* nq, const int* track, const int* sign);
int smdlt_download_charges(smdlt_ctx* ctx, int* track, int* old);

/* updateLattice(t) for every replica (S/main26.c:1382-1408, X/main14.c:1426-1450).
 * Host FP64; out (optional, per replica 8 doubles): L1x L1y L2x L2y L1xp L2xp L2yp delx. */
int smdlt_update_lattice(smdlt_ctx* ctx, uint64_t t, double* out);

/* The per-step remap on the device, FP64: KRBC transformCoords/inverseCoords
 * round trip (X/main14.c:69-71), updateLattice(t), transformCoords (X) and
 * applyPBC (S/main26.c:1409-1428, X/main14.c:1451-1461). */
int smdlt_remap(smdlt_ctx* ctx, uint64_t t);
/* applyPBC alone at the current lattice (positions already in the wrap frame). */
int smdlt_apply_pbc(smdlt_ctx* ctx);
/* transformCoords (inverse=0) / inverseCoords (inverse=1), X/main14.c:1514-1535. */
int smdlt_rotate(smdlt_ctx* ctx, int inverse);

/* getNID for a batch of pairs of one replica (S/main26.c:1429-1457,
 * X/main14.c:1462-1480): out[3*k] = r_i - r_j under the reference's minimum
 * image.  Computed on the device in FP64. */
int smdlt_min_image(smdlt_ctx* ctx, int replica, int npairs, const int* i, const int* j, double* out);
/* binij for a batch of pairs (S/main26.c:1518-1529) with the table geometry of
 * initLattice (bin_size = -g). */
int smdlt_bin_index(smdlt_ctx* ctx, int replica, double bin_size, int npairs, const int* i, const int* j,
                    int* bins);

/* checkVerlet + verletlist (S/main26.c:682-707, 858-869): rebuilds the LJ list
 * when any accumulated displacement exceeds (rv-rc)/2, or always if force != 0.
 * *rebuilt = number of replicas rebuilt.  The real-space mobility list is
 * refreshed by its own displacement test inside the same call. */
int smdlt_neighbors(smdlt_ctx* ctx, int force, int* rebuilt);
/* The reference's list of one replica: nlist[n], list[n*stride] ascending j>i. */
int smdlt_get_verlet(smdlt_ctx* ctx, int replica, int* nlist, int* list, int stride);

/* bondforce + LJforce + coulombForce (S/main26.c:541-681, 1891-1925).  Forces
 * stay on the device; fx/fy/fz may be NULL, else copied out (SoA doubles). */
int smdlt_forces(smdlt_ctx* ctx, double* fx, double* fy, double* fz);

/* Mobility product u = M v over all replicas (calcHI + the matvec of updateHI,
 * S/main26.c:1530-1806, 745-857).  v, u: AoS 3N doubles, host memory.
 * M is taken at the device positions in the frame the current lattice
 * describes, as calcHI runs after transformCoords/applyPBC (X/main14.c:77-79):
 * between steps a planar-extension replica's positions are in the lab frame,
 * so call smdlt_remap(t) first (a shear replica's frames coincide).  The
 * lists built by the last step or smdlt_neighbors are reused. */
int smdlt_mobility_apply(smdlt_ctx* ctx, const double* v, double* u);
/* Same with device pointers (no copies); stream-ordered on the context stream. */
int smdlt_mobility_apply_device(smdlt_ctx* ctx, const double* d_v, double* d_u);

/* Brownian displacement y = M^(1/2) z by Lanczos (replaces the TEA product of
 * S/main26.c:787-804).  iters[r], resid[r] per replica (may be NULL).  The stop
 * test |c_j - c_(j-1)| < tol |c_j| starts at iteration max(2, iters_prev / 2),
 * with iters_prev the replica's previous solve (0 on a new context). */
int smdlt_brownian(smdlt_ctx* ctx, const double* z, double* y, double tol, int maxit, int* iters,
                   double* resid);
/* The reference's TEA product y_i = C_i (.) (M_ii z_i + beta sum_j M_ij z_j) with
 * the mobility of this build (S/main26.c:787-804); beta per replica, C 3N. */
int smdlt_tea_apply(smdlt_ctx* ctx, const double* beta, const double* C, const double* z, double* y);

/* ewaldBin + CATEA (S/main26.c:975-1381, X/main14.c:1009-1420): one sample of
 * the conformation-averaged TEA parameters for each active replica, from the
 * device positions and the current lattice (the last smdlt_step, smdlt_remap
 * or smdlt_update_lattice), with the reference's Ewald sums (M_HI = 3.5, rce =
 * L2yp or 2 L2yp over the period, its self-term image sum and r <= 2 branch).
 * beta[nrep] (the TEA beta), C: AoS 3N (Ct), eps (optional) [nrep]: the pair
 * average eps / (9N^2 - 3N).  Entries of inactive replicas are left as they
 * are.  The strain-bin running averages of CATEA are the caller's; beta and C
 * feed smdlt_set_tea.  O(N^2) pair work per replica. */
int smdlt_tea_sample(smdlt_ctx* ctx, double* beta, double* C, double* eps);
/* Far-pair table of gridHI (S/gridHI.c:3-278, X/gridHI.c:3-270): the reference's
 * Ewald pair block (M_HI = 6, rce = L2yp or 2 L2yp over the period) at every bin
 * displacement of the table geometry of initLattice (bin_size = -g; the origin
 * bin at the reference's placeholder (0.6, 0.6, 0.6)), for nt_avg strain phases
 * t = nt*tp/nt_avg, from replica's L, flow, rate and dt.  table: nt_avg * 9 *
 * dims[0]*dims[1]*dims[2] doubles, [nt][bin][3x3 row-major], bin = (i*dims[1] +
 * j)*dims[2] + k; dims from smdlt_far_table_dims (sizes: the stretched bin
 * sizes, optional).  Flowing replicas only (SMDLT_E_ARG at equilibrium). */
int smdlt_far_table_dims(smdlt_ctx* ctx, int replica, double bin_size, int dims[3], double sizes[3]);
int smdlt_far_table(smdlt_ctx* ctx, int replica, int nt_avg, double bin_size, double* table);
/* Writes a table in the reference's mm/G text format (S/gridHI.c:247-263, read
 * by S/Initialization.c:497-520): per phase and bin, "i j k 1" then three rows
 * of three %.6e values. */
int smdlt_write_far_table(const char* path, const double* table, int nt_avg, const int dims[3]);

/* Virial stress summed over the steps since the last reset, per replica:
 * tau[3r..3r+2] = (xx, yy, xy) of -2 x_a F x_b over springs and LJ pairs plus the
 * reference's bending terms (the bondforce/LJforce accumulators tauxx, tauyy,
 * tauxy, S/main26.c:576-578,606-613,662-664, X/main14.c:583-619), lab frame.
 * Every step adds its forces' virial; smdlt_forces does not.  reset != 0 zeroes
 * the sums after the read, as printVisc does (S/main26.c:1866-1880, which writes
 * tauxy / printperiod; X/main14.c:1886-1897 writes tauxx and tauyy). */
int smdlt_stress(smdlt_ctx* ctx, double* tau, int reset);
/* Per-chain observables of the output step (calcExt, calcCOM, calcRg;
 * S/main26.c:74-93,925-962,2391-2415), chains of all replicas in order:
 * unwrapped chain p_{j+1} = p_j + getNID(j, j+1) from p_0 = 0, ext[k] = max - min
 * of p_x, com[3k..] = mean p, rg[6k..] = mean (p - com)(p - com) as xx, xy, xz,
 * yy, yz, zz.  Positions as the device holds them, in the frame of the current
 * lattice (call smdlt_remap(t) first for KRBC).  Any output may be NULL. */
int smdlt_chain_observables(smdlt_ctx* ctx, double* ext, double* com, double* rg);

/* The charge displacement log (dxt/dyt/dzt), summed since the last reset, per
 * charge in upload order: disp[3k..3k+2].  Every step adds the tracked bead's
 * displacement without the flow plus the flow relative to its chain's unwrapped
 * centre of mass (updateHI/updateFD, S/main26.c:815-838,2436-2446; X/main14.c:
 * 853-876); every accepted hop adds new bead minus old bead reduced by the cell
 * (chargeHop, S/main26.c:2221-2233).  calcCH prints these window sums
 * (S/main26.c:2452-2490); reset != 0 zeroes them after the read. */
int smdlt_charge_displacements(smdlt_ctx* ctx, double* disp, int reset);

/* TEA parameters for SMDLT_NOISE_TEA steps (the output of ewaldBin/CATEA,
 * S/main26.c:975-1381): beta per replica, C per bead component (3N AoS). */
int smdlt_set_tea(smdlt_ctx* ctx, const double* beta, const double* C);

/* One full timestep at step t (S/main26.c:65-159, X/main14.c:68-247, iteration
 * >= 1): remap, Verlet check, forces, noise, M F + M^(1/2) z (or TEA), position
 * update with the flow.  noise: NULL = device generator, else 3N host doubles. */
int smdlt_step(smdlt_ctx* ctx, uint64_t t, const double* noise);
/* nsteps steps from t0 with device noise.  use_graph != 0 captures one step in a
 * hipGraph and replays it (the lattice parameters are refreshed between replays). */
int smdlt_run(smdlt_ctx* ctx, uint64_t t0, int nsteps, int use_graph);
/* Per-replica Lanczos iteration counts of the last step (or brownian call). */
int smdlt_last_iters(const smdlt_ctx* ctx, int* iters);
/* Per-replica outcome of the last Lanczos solve: iters, status (1 converged,
 * 2 stopped at the iteration limit or on a breakdown, 0 not run) and the last
 * relative change |c_j - c_(j-1)| / |c_j|.  Any argument may be NULL. */
int smdlt_lanczos_status(smdlt_ctx* ctx, int* iters, int* status, double* change);

/* chargeHop (S/main26.c:1926-2390) for every replica at one MC step.  mode =
 * SMDLT_MC_*.  uniforms: per replica a host stream of n_uniforms values (layout
 * [replica][n_uniforms]) consumed in the reference's data-dependent order; NULL =
 * device generator.  consumed[r] = values used.  log (may be NULL): per replica
 * nq[r] records of 6 ints {k, from, to, kind(-1 skip,0 adj,1 adj-near,2 pop),
 * accepted, draws}, concatenated. */
int smdlt_mc_sweep(smdlt_ctx* ctx, int mode, const float* uniforms, int n_uniforms, int* consumed,
                   int* log);
/* smdlt_mc_sweep without allocating: no counterpart in the reference.  For
 * callers that cannot allocate device memory per call (tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments, results and errors as
 * smdlt_mc_sweep; the uniforms, consumed counts and reference-mode log are
 * staged in one library-owned device array per GPU (held until the process
 * ends) instead of per-call allocations.  Needs
 *   align256(4 nrep) + align256(24 nq_total) [reference mode only]
 *   + 4 n_uniforms nrep [host-fed uniforms only] <= SMDLT_MC_NOALLOC_MAX_BYTES
 * (align256 rounds up to a multiple of 256), else SMDLT_E_ARG.  All contexts
 * on one GPU share the array: calls that overlap in time on one GPU (from
 * several host threads) race; calls in sequence do not.  The MH modes' per-bead
 * neighbourhood lists live in the context with the capacity an earlier
 * smdlt_mc_sweep in an MH mode reserved; lists that would outgrow it (or no
 * such earlier sweep) give SMDLT_E_STATE, nothing is allocated. */
#define SMDLT_MC_NOALLOC_MAX_BYTES (8 << 20)
int smdlt_mc_sweep_noalloc(smdlt_ctx* ctx, int mode, const float* uniforms, int n_uniforms, int* consumed,
                           int* log);
/* Screened-Coulomb energy per replica (the double loop of chargeHop, S/main26.c:1955-1970). */
int smdlt_mc_energy(smdlt_ctx* ctx, double* energy);

/* Learned proposal for SMDLT_MC_MH_LEARNED: a 16-64-64-1 MLP with ReLU scoring
 * each hop candidate; weights are rounded to BF16 (nearest even), arithmetic is
 * FP32.  weights: FP32 host array of smdlt_mlp_num_params() values laid out
 * W1[64][16], b1[64], W2[64][64], b2[64], W3[64], b3.  No reference counterpart
 * (the reference's chargeHop proposes by a fixed rule, S/main26.c:1926-2390). */
int smdlt_mlp_num_params(void);
int smdlt_mlp_set_weights(smdlt_ctx* ctx, const float* weights);
/* Forward + backward of the mean binary cross-entropy (with logits) on a batch:
 * features [n][16] FP32, labels [n] in {0,1}; the weights are rounded to BF16 as
 * in the proposal, the layer products run on the matrix cores (BF16 in, FP32
 * accumulate).  grad gets d loss / d weights (FP32, smdlt_mlp_num_params()).
 * features = labels = NULL reuses the batch of the previous call (same n), which
 * stays on the device: a training loop uploads its data once. */
int smdlt_mlp_loss_grad(smdlt_ctx* ctx, const float* weights, const float* features, const float* labels,
                        int n, float* loss, float* grad);
/* The proposed moves of the last SMDLT_MC_MH_* sweep of one replica (one per
 * attempt, nq[replica] attempts): their 16 features, source and destination bead
 * (destination -1: no candidate) and whether they were accepted.  *n = count;
 * returns SMDLT_E_ARG if it exceeds cap.  Any output may be NULL. */
int smdlt_mc_features(smdlt_ctx* ctx, int replica, float* features, int* from, int* to, int* accepted, int cap,
                      int* n);
/* Whether SMDLT_MC_MH_* sweeps record the features of their proposed moves for
 * smdlt_mc_features (on by default; no reference counterpart).  Off saves that
 * work in uniform-proposal sweeps that no training reads; after such a sweep
 * smdlt_mc_features returns SMDLT_E_STATE when asked for features. */
int smdlt_mc_record_features(smdlt_ctx* ctx, int on);

/* Per-stage device times of the last step, milliseconds (HIP events on the
 * context stream when profiling is enabled): see smdlt_stage_name(). */
int smdlt_set_profiling(smdlt_ctx* ctx, int on);
int smdlt_stage_times(const smdlt_ctx* ctx, double* ms, int* counts, int max_stages);
const char* smdlt_stage_name(int stage);

/* Synchronise the context stream. */
int smdlt_sync(smdlt_ctx* ctx);

/* ------------------------------------------------------ run set-up (host)
 * The reference's random stream and initial state, on the host as in the
 * reference, so that a driver reproduces them from the same seed.  One stream
 * serves all draws of a run in the reference's order (initChains, then getNoise
 * and chargeHop every step). */
typedef struct smdlt_ran1 {
  int64_t idum, idum2, iy; /* ran1's seed word and its static state */
  int64_t iv[32];          /* shuffle table */
  int iset;                /* gasdev holds a second value */
  float gset;
} smdlt_ran1;
/* ran1 with the static initial state of a fresh process and *idum = seed
 * (830062920 shear, 981272199 planar extension; S/Initialization.c:438,
 * X/Initialization.c:426). */
void smdlt_ran1_seed(smdlt_ran1* s, int64_t seed);
/* ran1 (S/Initialization.c:601-634): the next uniform in (0, 1 - 1.2e-7]. */
float smdlt_ran1_next(smdlt_ran1* s);
void smdlt_ran1_fill(smdlt_ran1* s, int n, float* out);
/* getNoise (S/main26.c:515-540,708-713): n gasdev values (floats) as doubles;
 * a step's host noise is 3N of them. */
void smdlt_gasdev_fill(smdlt_ran1* s, int n, double* out);
/* initChains for a new run (S/main26.c:346-470): nc self-avoiding chains of nb
 * beads with bonds 2.05 in the cube [-L/2, L/2)^3, positions chain by chain.
 * SMDLT_E_GEOMETRY if the box is too full. */
int smdlt_init_chains(int nb, int nc, double L, smdlt_ran1* s, double* rx, double* ry, double* rz);
/* The charge placement of initChains (S/main26.c:485-500): ncharges distinct
 * beads of n, in draw order (Charge_track). */
int smdlt_init_charges(int n, int ncharges, smdlt_ran1* s, int* track);
/* The restart reader of initChains (S/main26.c:264-275): the last complete
 * frame of "N", "t", N lines "A x y z" (printRestart's format).  Returns n, or
 * SMDLT_E_ARG (no file, no frame, or a frame of another size); *t may be NULL. */
int smdlt_read_restart(const char* path, int n, double* rx, double* ry, double* rz, uint64_t* t);

/* ------------------------------------------- TEA averages and hand-off (host)
 * Iteration 0 of the reference samples the TEA parameters (smdlt_tea_sample)
 * and averages them per strain bin; iteration 1 steps with the bin averages
 * (SMDLT_NOISE_TEA).  One averaging object per replica. */
typedef struct smdlt_tea_avg smdlt_tea_avg;
/* The running sums of CATEA (S/main26.c:1317-1381, S/Initialization.c:401-428):
 * sbins + 1 strain bins of width strain_interval (the last one collects every
 * larger strain), per bin the sum of beta, of C by index within a chain
 * (3 nb values) and the sample count. */
int smdlt_tea_avg_create(int nb, int nc, int sbins, double strain_interval, smdlt_tea_avg** out);
void smdlt_tea_avg_destroy(smdlt_tea_avg* avg);
/* CATEA's bookkeeping for one sample at step t and strain (S/main26.c:1346-1366):
 * bin = floor(strain / strain_interval) capped at sbins; beta and C[3 nb nc]
 * (one replica's smdlt_tea_sample output) added to the bin.  beta_path (may be
 * NULL) gets CATEA's decomp/B record, "t beta" then "i Cx Cy Cz" per bead
 * (S/main26.c:1371-1376).  Returns the bin. */
int smdlt_tea_avg_add(smdlt_tea_avg* avg, uint64_t t, double strain, double beta, const double* C,
                      const char* beta_path);
/* printTEA (S/main26.c:1807-1826): writes the decomp/DC file: "t", then per bin
 * "bin count mean-beta" and nb rows "j Cx Cy Cz" of mean C (zeros for an empty bin). */
int smdlt_tea_avg_write(const smdlt_tea_avg* avg, uint64_t t, const char* path);
/* resetAverage (S/main26.c:1827-1839): the per-bin means for iteration 1,
 * beta[sbins + 1] and C[(sbins + 1) * 3 nb] (an empty bin gives 0/0, as in the
 * reference), counts[sbins + 1]; the sums restart from zero.  Any output may be NULL. */
int smdlt_tea_avg_finish(smdlt_tea_avg* avg, double* beta, double* C, int* counts);
/* The decomp/DC reader of a restarted iteration (S/Initialization.c:522-540). */
int smdlt_tea_read(const char* path, int nb, int sbins, double* beta, double* C, uint64_t* t);
/* updateHI's choice of coefficients at a strain (S/main26.c:757-764): the bin's
 * beta and C[3 nb nc] (C of bead i is the bin's entry i mod 3 nb), ready for
 * smdlt_set_tea.  Returns the bin. */
int smdlt_tea_coefficients(int nb, int nc, int sbins, double strain_interval, double strain, const double* beta_bins,
                           const double* C_bins, double* beta, double* C);

/* ------------------------------------------------ MSD post-processing (host)
 * calcMSD (S/main26.c:2494-2637): reads the calcCH file (ncharges lines "t j old
 * track dx dy dz" for each of the tsteps + 1 output steps, tsteps = tmax /
 * printprops), and writes prop/MSD rows "tau,MSD,MSDx,MSDy,MSDz" for shifts
 * 1 .. tsteps - msd_start, tau = shift dt printprops, averaged over the first
 * min(25, ncharges) charges with the reference's index layout (series q > 0
 * read from q tsteps, one entry before the slot its sort put them in).
 * out_path and out (rows x 5 doubles, at most cap rows) may be NULL.  Returns
 * the row count. */
int smdlt_msd(const char* ch_path, int ncharges, int tsteps, int msd_start, double dt, int printprops,
              const char* out_path, double* out, int cap);
/* calcMSDcom (X/main14.c:3096-3214): one series of lines "t dx dy dz", shifts
 * 1 .. tsteps - 1, windows from the second entry. */
int smdlt_msd_com(const char* com_path, int tsteps, double dt, int printprops, const char* out_path, double* out,
                  int cap);

#ifdef __cplusplus
}
#endif
#endif
