/*
 * smdlt_oracle.h -- CPU FP64 restatement of the reference algorithm for the
 * Brownian-dynamics timestep hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product (libsmdlt_hip.so) never links, loads or calls it.
 *
 * Every function cites the reference lines it restates (S/ = shear/src/,
 * X/ = extensional/src/).  State that the reference keeps in globals is passed
 * explicitly here.  Arithmetic is written in the reference's operation order
 * and the file is compiled with -ffp-contract=off so that integer outputs
 * (lists, bins, MC index sequences) can be compared bit-exactly.
 */
#ifndef SMDLT_ORACLE_H
#define SMDLT_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- RNG: ran1 + gasdev (S/Initialization.c:601-634, S/main26.c:515-540) */
typedef struct {
  int64_t idum, idum2, iy;
  int64_t iv[32];
  int iset;
  float gset;
} orc_rng_t;
void  orc_rng_seed(orc_rng_t* s, int64_t seed);
float orc_ran1(orc_rng_t* s);
float orc_gasdev(orc_rng_t* s);
/* fill out[n] with successive ran1 values */
void  orc_ran1_fill(orc_rng_t* s, int n, float* out);
void  orc_gasdev_fill(orc_rng_t* s, int n, double* out);
int   orc_init_chains(orc_rng_t* s, int Nb, int Nc, double L, double* rx, double* ry, double* rz);
void  orc_init_charges(orc_rng_t* s, int N, int Ncharges, int* track);

/* ---------------- Lattice (S/main26.c:1382-1457, X/main14.c:1426-1535) */
typedef struct {
  int flow;                 /* 0 equilibrium (cubic), 1 shear (Lees-Edwards), 2 extension (KRBC) */
  double L, edot, dt;
  int tp;                   /* lattice reset period in steps */
  double point0[3][2], point[3][2];
  double L1[2], L2[2];
  double L1xp, L1yp, L2xp, L2yp, detL, L2pmag;
  int treal;                /* t % tp */
  double delx;              /* shear: edot*treal*dt (the x shift quirk) */
  double theta;             /* atan(L1[1]/L1[0]) */
  double ct, st;            /* cos(theta), sin(theta) */
  int pbc_compat;           /* 1: reference delx shift (shear), 0: plain lattice wrap */
} orc_lattice_t;
void orc_lattice_init(orc_lattice_t* lat, int flow, double L, double edot, double dt, int pbc_compat);
void orc_lattice_update(orc_lattice_t* lat, uint64_t t);
void orc_apply_pbc(const orc_lattice_t* lat, int n, double* rx, double* ry, double* rz);
void orc_rotate(const orc_lattice_t* lat, int n, double* rx, double* ry, int inverse);
void orc_nid(const orc_lattice_t* lat, const double* rx, const double* ry, const double* rz,
             int i, int j, double out[3]);
void orc_nid_bin(const orc_lattice_t* lat, const double* rx, const double* ry, const double* rz,
                 int i, int j, double out[3]);

/* far-field table binning (S/Initialization.c:217-298, S/main26.c:1518-1529) */
typedef struct {
  double bin_size_x, bin_size_y, bin_size_z;
  int nbx, nby, nbz, nhx, nhy, nhz, num_bins;
} orc_bins_t;
void orc_bins_init(orc_bins_t* b, const orc_lattice_t* lat_at_init, double bin_size);
int  orc_binij(const orc_bins_t* b, const orc_lattice_t* lat, const double* rx, const double* ry,
               const double* rz, int i, int j);

/* ---------------- Neighbour list (S/main26.c:682-707, 858-869) */
void orc_verlet(const orc_lattice_t* lat, int n, const double* rx, const double* ry, const double* rz,
                double rv, int* nlist, int* list, int stride);
int  orc_check_verlet(int n, const double* dxv, const double* dyv, const double* dzv, double rnew);

/* ---------------- Conservative forces (S/main26.c:541-681, 1891-1925; X/main14.c LJ) */
int  orc_bond_force(const orc_lattice_t* lat, int nb, int nc, const double* rx, const double* ry,
                    const double* rz, int spring_fene, double kappas, double qmax, double kappab,
                    double* fx, double* fy, double* fz, double tau[3]);
void orc_chain_obs(const orc_lattice_t* lat, int nb, int nc, const double* rx, const double* ry,
                   const double* rz, double* ext, double* com, double* rg);
void orc_lj_force(const orc_lattice_t* lat, int n, const double* rx, const double* ry, const double* rz,
                  const int* nlist, const int* list, int stride, double r2cut, double epsilon,
                  int ext_form, double* fx, double* fy, double* fz, double tau[3]);
void orc_coulomb_force(const orc_lattice_t* lat, int n, const double* rx, const double* ry,
                       const double* rz, const int* q, double lambda_b, double lambda_d,
                       double* fx, double* fy, double* fz);
double orc_coulomb_energy(const orc_lattice_t* lat, int n, const double* rx, const double* ry,
                          const double* rz, const int* q, double lambda_b, double lambda_d);

/* ---------------- Exact Ewald-summed RPY (Beenakker; a=1, kT=1, unit self-mobility)
 * General deformed lattice a1=(a1x,a1y,0), a2=(a2x,a2y,0), a3=(0,0,Lz).
 * pos is AoS 3n.  M is the full 3n x 3n matrix, row-major.
 * overlap_mode 0: exact overlap-corrected RPY for r<2; 1: the reference's r<=2
 * branch (S/main26.c:1734-1739). */
typedef struct { double a1x, a1y, a2x, a2y, Lz; } orc_cell_t;
void orc_rpy_real_coeffs(double r, double alpha, int overlap_mode, double* t1, double* t2);
void orc_rpy_ewald_dense(const orc_cell_t* c, int n, const double* pos, double alpha, double rcut,
                         double kcut, int overlap_mode, double* M);
/* u = M v with M of orc_rpy_ewald_dense, without forming M (structure factors
 * for the reciprocal part): for checks at 10^4 - 10^5 beads. */
void orc_rpy_ewald_matvec(const orc_cell_t* c, int n, const double* pos, double alpha, double rcut, double kcut,
                          int overlap_mode, const double* v, double* u);
/* free-space RPY (no periodic images), for the L -> infinity closed form */
void orc_rpy_free_dense(int n, const double* pos, double* M);
double orc_ewald_selfconst(double alpha);

/* reference calcHI restated with all its quirks (S/main26.c:1530-1806, X/main14.c:1548-1824).
 * Positions in the reference's storage convention for the variant.  table may be
 * NULL: then far pairs (>= 12a) are evaluated by the same near-pair formula and
 * *far_count reports how many pairs that affected. M: upper blocks j>=i written,
 * lower blocks mirrored so the result is a full symmetric matrix. */
void orc_calchi_ref(const orc_lattice_t* lat, int n, const double* rx, const double* ry,
                    const double* rz, uint64_t t, int nt_avg, const orc_bins_t* bins,
                    const double* table /* nt_avg*9*num_bins or NULL */, double box_volume,
                    double* M, int* far_count, int* kcount_out);
void orc_ewald_bin(const orc_lattice_t* lat, int N, int Nb, const double* rx, const double* ry, const double* rz,
                   double box_volume, double M_HI, double* eps, double* rowsum, double Dtr[3], int* kcount_out);
double orc_catea(int N, double eps_sum, const double* rowsum, double* C);
void orc_far_table(const orc_lattice_t* lat0, int nt_avg, const orc_bins_t* b, double box_volume, double* table);

/* ---------------- Brownian displacement helpers */
/* TEA product (S/main26.c:787-804): y_i = C_i (.) (M_ii R_i + beta sum_{j!=i} M_ij R_j),
 * with the reference's beta factor on the off-diagonal elements of M_ii. */
void orc_tea_apply(int n, const double* M, double beta, const double* C, const double* R, double* y);
/* symmetric square root by cyclic Jacobi: y = A^{1/2} z (A is n3 x n3, SPD) */
int  orc_sym_sqrt_apply(int n3, const double* A, const double* z, double* y, double* eigmin,
                        double* eigmax);
/* Lanczos y ~ A^{1/2} z with the coefficient-change stopping rule */
int  orc_lanczos_sqrt(int n3, const double* A, const double* z, double tol, int maxit, double* y,
                      int* iters);
void orc_matvec(int n3, const double* A, const double* x, double* y);

/* ---------------- Charge-hopping Monte Carlo (S/main26.c:1926-2390) */
typedef struct {
  int n, nb, nc, nq;
  int* track;      /* Charge_track[nq] */
  int* indices;    /* Charge_indices[nq] */
  int* old;        /* Charge_old[nq] */
  int* ivalues;    /* i_values[n] */
  int* charge;     /* Charge[n] scratch (0/1) */
} orc_mc_state_t;
/* one chargeHop call at a step with t % MStep == 0.  Uniforms are consumed from
 * u[*cursor...]; log (may be NULL) gets 6 ints per charge k:
 * {k, cnumber, cnumber2, test(0 adj,1 adj-near,2 pop,-1 skipped), accepted, draws}. */
int orc_charge_hop(orc_mc_state_t* s, const orc_lattice_t* lat, const double* rx, const double* ry,
                   const double* rz, double lambda_b, double lambda_d, double barrier,
                   double barrier2, const float* u, int nu, int* cursor, int* log,
                   double* e_initial_out);

/* ---------------- Metropolis-Hastings charge hopping (no reference counterpart: an
 * exactly detailed-balanced alternative to chargeHop's proposal, S/main26.c:1926-2390,
 * with the same screened-Coulomb energy and barriers).
 * A move takes charge k (sign qk) from bead x to an uncharged bead y with
 * |getNID(x,y)| < hop_radius, drawn from q(x->.) (uniform, or learned: a
 * (1-floor)/floor mixture of the normalised predicted acceptances sigmoid(MLP
 * score), a Barker-type locally balanced weight, and uniform).  With
 * rho = exp(-dE) q(y->x)/q(x->y) and a = exp(-b) (b = barrier for chain
 * neighbours, barrier2 otherwise) the move is accepted with probability
 * A = a sqrt(rho) / max(1, a sqrt(rho), a / sqrt(rho)), so A(x->y)/A(y->x) = rho
 * (detailed balance for exp(-E)) and A = exp(-b - dE/2), the reference's rule,
 * whenever |ln rho| <= 2b under a symmetric proposal.  Three uniforms per attempt:
 * u0 picks k = min(floor(u0 nq), nq-1), u1 picks y by inverse CDF over the
 * candidates in ascending bead order, u2 is the acceptance dice (accept if u2 < A). */
typedef struct {
  int n, nb;                /* beads, beads per chain */
  double hop_radius, feature_radius, floor_eps;
  double lambda_b, lambda_d, barrier, barrier2;
} orc_mh_params_t;
/* round a float to bfloat16 (round to nearest even), returned as float */
float orc_bf16_round(float x);
/* The 16 proposal features of moving the charge at bead src to bead dst.  qbead[n]:
 * charge sign per bead (0 = none) with the moving charge at src; clr / set (or -1):
 * a bead whose charge is taken as 0 / a bead taken to carry the moving charge
 * (the state after a move, for the reverse proposal).  ncand: candidates of src.
 * The distances of src and dst to the other beads are the norms of differences
 * of their getNID vectors relative to bead anchor (the attempt's start bead). */
void orc_mh_features(const orc_lattice_t* lat, const orc_mh_params_t* p, const double* rx, const double* ry,
                     const double* rz, const int* qbead, int src, int dst, int qk, int clr, int set, int ncand,
                     int anchor, double f[16]);
/* MLP 16-64-64-1 with ReLU, FP64 arithmetic on the given (already rounded) weights
 * laid out W1[64][16], b1[64], W2[64][64], b2[64], W3[64], b3. */
double orc_mlp_forward(const float* w, const double f[16]);
/* nattempt attempts; w = NULL: uniform proposal.  u: 3 per attempt.  log: 6 ints per
 * attempt {k, x, y, kind (0 chain neighbour, 2 other, -1 no candidate), accepted, 3};
 * feats (may be NULL): 16 features of each proposed move (zeros when skipped).
 * track/sign: charges; qbead: per-bead sign, kept consistent.  Returns accepted count. */
int orc_mh_sweep(const orc_lattice_t* lat, const orc_mh_params_t* p, const double* rx, const double* ry,
                 const double* rz, int nq, int* track, const int* sign, int* qbead, const float* w,
                 const float* u, int nattempt, int* log, float* feats);

/* ---------------- The reference's HI timestep, restated for CPU timing.
 * Dense Ewald M rebuilt every srpy steps, TEA matvec, forces, MC every mstep.
 * OpenMP inside; returns elapsed seconds. */
typedef struct {
  int nb, nc, flow;
  double L, edot, dt, kappas, qmax, epsilon, lambda_b, lambda_d, barrier, barrier2;
  int spring_fene, nq, mstep, srpy, nt_avg;
  double beta;     /* TEA beta */
  int64_t seed;
} orc_ref_run_t;
double orc_ref_run(const orc_ref_run_t* p, double* rx, double* ry, double* rz, int nsteps,
                   int nthreads, int* status, double* t_hi);

int orc_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
