// Host side of the C ABI declared in include/smdlt_hip.h: context and device
// memory, the flowing-box lattice (updateLattice restated in FP64 with the
// reference's expressions), the mobility settings (autotuner), the step
// sequence and the stage timers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <functional>
#include <mutex>
#include <vector>

#include "common.h"
#include "smdlt_hip.h"

#define HCK(x)                                   \
  do {                                           \
    hipError_t e_ = (x);                         \
    if (e_ != hipSuccess) {                      \
      last_hip_error() = hipGetErrorString(e_);  \
      return SMDLT_E_HIP;                        \
    }                                            \
  } while (0)

static std::string& last_hip_error() {
  static thread_local std::string s;
  return s;
}

static int resolve_scratch(int device);  // device address of smdlt_mc_sweep_noalloc's buffer

static const double kEP = 0.9624;  // lattice reset strain (S/Parameters.h:37)

// ---------------------------------------------------------------- lattice
// initLattice / updateLattice of the reference (S/Initialization.c:217-298,
// S/main26.c:1382-1408; X/Initialization.c:236-287, X/main14.c:1426-1450).
struct HostLattice {
  int flow = 0, tp = INT_MAX, pbc_compat = 1, treal = 0;
  double L = 0, edot = 0, dt = 0;
  double point0[3][2], point[3][2], L1[2], L2[2];
  double L1xp = 0, L1yp = 0, L2xp = 0, L2yp = 0, detL = 0, delx = 0, theta = 0, ct = 1, st = 0;

  void init(int flow_, double L_, double edot_, double dt_, int compat) {
    flow = flow_;
    L = L_;
    edot = edot_;
    dt = dt_;
    pbc_compat = compat;
    tp = (flow == 0 || edot == 0.0) ? INT_MAX : (int)(kEP / edot / dt);
    double theta0 = (flow == 2) ? 31.7 : 0.0;
    double th = 90.0 - (45.0 + theta0);
    th *= M_PI / 180;
    point[0][0] = -sqrt(2) / 2 * L * cos(th);
    point[0][1] = sqrt(2) / 2 * L * sin(th);
    point[1][0] = -sqrt(2) / 2 * L * sin(th);
    point[1][1] = -sqrt(2) / 2 * L * cos(th);
    point[2][0] = sqrt(2) / 2 * L * cos(th);
    point[2][1] = -sqrt(2) / 2 * L * sin(th);
    memcpy(point0, point, sizeof(point));
    for (int i = 0; i < 2; ++i) {
      L1[i] = point[2][i] - point[1][i];
      L2[i] = point[0][i] - point[1][i];
    }
    update(0);
  }
  void update(uint64_t t) {
    treal = (tp == INT_MAX) ? 0 : (int)(t % (uint64_t)tp);
    double flowrate = edot;
    if (flow == 2) {
      for (int i = 0; i < 3; ++i) {
        point[i][0] = point0[i][0] * exp(flowrate * treal * dt);
        point[i][1] = point0[i][1] * exp(-flowrate * treal * dt);
      }
      for (int i = 0; i < 2; ++i) {
        L1[i] = point[2][i] - point[1][i];
        L2[i] = point[0][i] - point[1][i];
      }
      double th = atan(L1[1] / L1[0]);
      L1xp = L1[0] * cos(th) + L1[1] * sin(th);
      L2xp = L2[0] * cos(th) + L2[1] * sin(th);
      L1yp = -L1[0] * sin(th) + L1[1] * cos(th);
      L2yp = -L2[0] * sin(th) + L2[1] * cos(th);
      detL = L * (-L1yp * L2xp + L1xp * L2yp);
      delx = 0.0;
    } else {
      point[0][0] = point0[0][0] + point[0][1] * flowrate * treal * dt;
      point[1][0] = point0[1][0] + point[1][1] * flowrate * treal * dt;
      point[2][0] = point0[2][0] + point[2][1] * flowrate * treal * dt;
      for (int i = 0; i < 2; ++i) {
        L1[i] = point[2][i] - point[1][i];
        L2[i] = point[0][i] - point[1][i];
      }
      L1xp = L1[0];
      L2xp = L2[0];
      L1yp = L1[1];
      L2yp = L2[1];
      detL = L2yp * L1xp * L;
      delx = flowrate * treal * dt;
    }
    theta = atan(L1[1] / L1[0]);
    ct = cos(theta);
    st = sin(theta);
  }
};

// Lagrange-Gauss reduction of the in-plane basis (a1, a2).
static void reduce_basis(double a1x, double a1y, double a2x, double a2y, double b[4], double dual[4], double* wmin) {
  double u[2] = {a1x, a1y}, v[2] = {a2x, a2y};
  for (int it = 0; it < 64; ++it) {
    double uu = u[0] * u[0] + u[1] * u[1], vv = v[0] * v[0] + v[1] * v[1];
    if (vv < uu) {
      std::swap(u, v);
      std::swap(uu, vv);
    }
    double mu = std::round((u[0] * v[0] + u[1] * v[1]) / uu);
    if (mu == 0.0) break;
    v[0] -= mu * u[0];
    v[1] -= mu * u[1];
  }
  b[0] = u[0];
  b[1] = u[1];
  b[2] = v[0];
  b[3] = v[1];
  double det = u[0] * v[1] - u[1] * v[0];
  dual[0] = v[1] / det;
  dual[1] = -v[0] / det;
  dual[2] = -u[1] / det;
  dual[3] = u[0] / det;
  double area = fabs(det);
  double w1 = area / sqrt(v[0] * v[0] + v[1] * v[1]), w2 = area / sqrt(u[0] * u[0] + u[1] * u[1]);
  *wmin = std::min(w1, w2);
}

static void reduced_widths(const HostLattice& h, double w[2]) {
  double b[4], d[4], wm;
  reduce_basis(h.L1xp, 0.0, h.L2xp, h.L2yp, b, d, &wm);
  double area = fabs(h.L1xp * h.L2yp);
  w[0] = area / sqrt(b[2] * b[2] + b[3] * b[3]);
  w[1] = area / sqrt(b[0] * b[0] + b[1] * b[1]);
}

// ---------------------------------------------------------------- mobility settings
static double ewald_selfconst(double a) {
  double rtpi = sqrt(M_PI);
  return 1 - 6 / rtpi * a + 40 / (3 * rtpi) * a * a * a;
}

// magnitude of the Beenakker real-space coefficients at r (far form)
static double real_tail(double r, double a) {
  double rr = r * r, a2 = a * a, a3 = a2 * a, a5 = a3 * a2, a7 = a5 * a2;
  double C1 = 0.75 / r + 0.5 / (r * rr), C3 = 0.75 / r - 1.5 / (r * rr);
  double C2 = 4 * a7 * rr * rr + 3 * a3 * rr - 20 * a5 * rr - 4.5 * a + 14 * a3 + a / rr;
  double C4 = -4 * a7 * rr * rr - 3 * a3 * rr + 16 * a5 * rr + 1.5 * a - 2 * a3 - 3 * a / rr;
  double ec = erfc(a * r), g = exp(-a2 * rr) / sqrt(M_PI);
  return fabs(C1 * ec + C2 * g) + fabs(C3 * ec + C4 * g);
}

// alpha such that the real-space coefficient at rc is tol * (density factor)
static double choose_alpha(double rc, double tol) {
  double lo = 0.01, hi = 5.0;
  for (int i = 0; i < 100; ++i) {
    double mid = 0.5 * (lo + hi);
    if (real_tail(rc, mid) > tol)
      lo = mid;
    else
      hi = mid;
  }
  return hi;
}

// k at which the reciprocal kernel envelope (1+q+2q^2)e^{-q} 6pi/k^2 drops to tol
static double choose_kmax(double alpha, double tol) {
  double k = 0.1;
  for (; k < 100.0; k *= 1.01) {
    double q = k * k / (4 * alpha * alpha);
    double env = (1 + q + 2 * q * q) * exp(-q) * 6 * M_PI / (k * k) * fabs(1 - k * k / 3);
    if (env < tol && q > 1.0) break;
  }
  return k;
}

// FFT lengths: multiples of 8 (grid bricks) of the form 2^a 3^b 5^c, 16..KMAX
static bool good_fft_len(int k) {
  if (k < 16 || k > KMAX || k % 8) return false;
  for (int f : {2, 3, 5})
    while (k % f == 0) k /= f;
  return k == 1;
}
static int next_fft_len(int x) {
  int k = std::max(16, x);
  while (k <= KMAX && !good_fft_len(k)) ++k;
  return k;  // > KMAX when x is too large (rejected by the caller)
}

// B-spline modulus |b(m)|^2 = 1/|sum_{k=0}^{p-2} M_p(k+1) e^{2 pi i m k/K}|^2 (Essmann et al. 1995)
static void bspline_moduli(int K, int p, std::vector<double>& out) {
  std::vector<double> W(MAX_ORDER, 0.0);
  // M_p at integer points 1..p-1: weights W[j] = M_p(0 + j) at x=0
  W[0] = 1.0;
  for (int k = 2; k <= p; ++k) {
    double div = 1.0 / (k - 1);
    W[k - 1] = div * W[k - 2];
    for (int j = k - 2; j >= 1; --j) W[j] = div * (j * W[j] + (k - j) * W[j - 1]);
    W[0] = 0.0;
  }
  out.assign(K, 0.0);
  for (int m = 0; m < K; ++m) {
    double re = 0, im = 0;
    for (int k = 0; k <= p - 2; ++k) {
      double ang = 2 * M_PI * m * k / K;
      re += W[k + 1] * cos(ang);
      im += W[k + 1] * sin(ang);
    }
    double d = re * re + im * im;
    out[m] = d > 1e-10 ? 1.0 / d : 0.0;
  }
}

// ---------------------------------------------------------------- context
enum Stage {
  ST_REMAP, ST_NLIST, ST_FORCES, ST_NOISE, ST_MOB, ST_LANCZOS, ST_UPDATE, ST_MC,
  ST_K_ZERO, ST_K_SPREAD, ST_K_FFTX, ST_K_FFTY, ST_K_FFTZ_GREEN, ST_K_IFFTY, ST_K_IFFTX, ST_K_FINAL, ST_COUNT
};
static const char* kStageNames[ST_COUNT] = {
    "remap", "neighbors", "forces", "noise", "mobility_apply", "lanczos_other", "update", "mc",
    "k_bead_prep", "k_spread", "k_fft_x(fwd)", "k_fft_y(fwd)", "k_fft_z_green", "k_fft_y(inv)",
    "k_fft_x(inv)", "k_mob_final"};

struct Timer {
  hipEvent_t a, b;
  int stage;          // -1: events only, nothing accumulated
  bool own_a = true;  // return a to the pool after reading (shared events are returned once)
  int calls = 1;      // calls counted (a pass split over FFT chunks counts once)
};

struct smdlt_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int nrep = 0, prec = 0, nst = 0, nwaves = 0;
  std::vector<smdlt_params> P;
  std::vector<HostLattice> lat;
  std::vector<RepDev> rh;
  std::vector<double> ct_prev, st_prev;
  std::vector<int> max_w_cells;  // cell capacity per replica
  RepDev* d_reps = nullptr;
  int* d_slot_rep = nullptr;
  uint64_t* d_seeds = nullptr;
  // bead state
  double *rx = nullptr, *ry = nullptr, *rz = nullptr, *dxv = nullptr, *dyv = nullptr, *dzv = nullptr;
  double *dsp = nullptr, *f = nullptr, *z = nullptr, *mf = nullptr, *bz = nullptr, *tmp = nullptr, *tmp2 = nullptr;
  bool have_pos = false;
  // lists
  int cap_rpy = 64, cap_lj = 64;
  double nbr_est = 0;
  int *nrpy = nullptr, *lrpy = nullptr, *nlj = nullptr, *llj = nullptr, *overflow = nullptr;
  int *cell_of = nullptr, *cell_cnt = nullptr, *cell_start = nullptr, *cell_fill = nullptr, *cell_beads = nullptr,
      *cell_tmp = nullptr;
  int ncells = 0;
  int* rep_flags = nullptr;
  double *out2 = nullptr, *aff = nullptr;
  bool lists_built = false;
  // grid
  void* grid = nullptr;
  long long grid_total = 0;
  void* bsp = nullptr;
  void* tw = nullptr;
  FftTask *tx = nullptr, *ty = nullptr, *tz = nullptr, *txy = nullptr;
  int ntx = 0, nty = 0, ntz = 0, ntxy = 0;
  int ztile = FFT_TILE_Z;  // complex elements per component of the z tasks
  BrickTask* bricks = nullptr;
  int nbricks = 0;
  std::vector<int> bk0;  // per-replica brick task ranges (size nrep+1)
  int *bbase = nullptr, *brick_cnt = nullptr, *brick_start = nullptr, *brick_fill = nullptr, *brick_beads = nullptr;
  void* bw = nullptr;
  uint64_t pos_version = 1, prep_version = 0;
  long long prep_range = -1;
  std::vector<int> tx0, ty0, tz0, txy0;  // per-replica task ranges (size nrep+1)
  std::vector<int> fch_rep;  // first replicas of the FFT chunks after the first (FFT_CHUNK_BYTES)
  std::vector<int> fch_buf;  // the chunk table of the last mob_args call (MobArgs::fch)
  // lanczos
  int maxit = 30;
  double *V = nullptr, *W = nullptr, *partial = nullptr, *partial2 = nullptr, *scal = nullptr, *bdiv = nullptr;
  int *done = nullptr, *iters = nullptr;
  std::vector<int> last_iters;
  // charges
  std::vector<int> nq, q0;
  int nq_total = 0, max_nq = 0;
  int *qtrack = nullptr, *qsign = nullptr, *qold = nullptr, *qidx = nullptr, *mc_err = nullptr;
  unsigned char* occ = nullptr;
  int *mc_start = nullptr, *mc_jf = nullptr;  // reference sweep: start beads, j* per charge
  double* mc_de = nullptr;  // reference sweep: candidate energy changes [nq_total][3]; smdlt_mc_energy: row sums
  void* mc_state = nullptr;                    // reference sweep: per-replica cursor, draws, E - E_initial
  int *mh_cnt = nullptr, *mh_off = nullptr;    // MH sweeps: neighbourhood list counts and offsets [nst + 1]
  int* mh_list = nullptr;                      // MH sweeps: neighbourhood lists
  long long mh_list_cap = 0;
  int *mh_hcnt = nullptr, *mh_hoff = nullptr, *mh_hlist = nullptr;  // MH sweeps: hop lists
  long long mh_hlist_cap = 0;
  double* mh_V = nullptr;                      // MH sweeps: charge potential per bead [nst]
  void* mh_moves = nullptr;                    // MH sweeps: a round's accepted moves [nrep][64] int4
  int *mh_nmoves = nullptr, *mh_done = nullptr, *mh_stop = nullptr;  // [nrep]
  long long occ_stride = 0;
  int* log_off = nullptr;
  float4* p32 = nullptr;      // mixed precision: FP32 positions and scaled input vector
  float4* v32 = nullptr;
  int *bcand_cnt = nullptr, *bcand_start = nullptr, *bcand = nullptr;  // mixed precision: spread candidates per brick
  float4* rtab = nullptr;     // mixed precision: real-space coefficient tables
  // bead clusters and the cluster-pair list (mixed precision)
  int *cl_first = nullptr, *cl_cnt = nullptr, *slot_cl = nullptr, *lcl = nullptr, *nclp = nullptr;
  double* adot_cl = nullptr;  // per-cluster partial of the Lanczos alpha, written by the cluster pass
  int ncl_total = 0, capc = 0, cl_words = 0;
  float* mlp_w = nullptr;     // BF16-rounded proposal network weights (FP32 storage)
  float* mh_feats = nullptr;  // features of each MH attempt of the last sweep [nq_total][16]
  int* mh_log = nullptr;      // its log [nq_total][6]
  bool mh_swept = false;      // an MH sweep has filled mh_log
  bool mh_record = true;      // MH sweeps record the features of each proposed move
  bool mh_feats_valid = false;  // mh_feats holds the last MH sweep's features
  float* mlp_ws = nullptr;    // training: weights, batch, per-block partial gradients
  float* mlp_x = nullptr;
  float* mlp_y = nullptr;
  float* mlp_part = nullptr;
  float* mlp_out = nullptr;
  int mlp_cap = 0;
  int mlp_n = 0;              // size of the batch resident on the device
  uint64_t mc_sweeps = 0;
  int* d_err = nullptr;
  // profiling
  bool prof = false;
  std::vector<Timer> timers;
  std::vector<hipEvent_t> pool;
  double stage_ms[ST_COUNT] = {0};
  int stage_cnt[ST_COUNT] = {0};
  // graph
  // one step = head graph (remap .. M.F, Lanczos iterations 1..LZ_CHUNK), then
  // chunk graphs of LZ_CHUNK further iterations launched while any replica is
  // unconverged, then the tail graph (combination, position update)
  hipGraphExec_t gexec = nullptr;
  hipGraphExec_t gtail = nullptr;
  std::vector<hipGraphExec_t> gchunks;
  int* h_flags = nullptr;  // pinned: done[nrep], iters[nrep], overflow
  double *tea_beta = nullptr, *tea_C = nullptr, *tea_diag = nullptr;  // SMDLT_NOISE_TEA steps
  double* saved = nullptr; // positions, displacement trackers and stress sums at the start of a graph step
  // observables: per-bead virial terms, per-replica stress sums (xx, yy, xy),
  // first chain of each replica, per-chain extension/com/gyration output
  double *vir = nullptr, *tau = nullptr, *obs = nullptr;
  // charge displacement log: per-charge sums, their copy at the start of a graph
  // step, and each bead's offset from its chain's centre of mass (step start)
  double *qdisp = nullptr, *qdisp_saved = nullptr, *qrel = nullptr;
  int* d_chain0 = nullptr;
  std::vector<int> chain0;
  int nchains = 0, max_chains = 0;
  bool force_rebuild = false;
  uint64_t lat_t = UINT64_MAX;  // step of the last lattice set on the device
  int expect_iters = 0;    // largest iteration count of the previous graph step
  bool graph_ready = false;
  int cur_rep_begin = 0, cur_rep_end = 0;
};

static void destroy_graphs(smdlt_ctx* c) {
  if (c->gexec) hipGraphExecDestroy(c->gexec);
  if (c->gtail) hipGraphExecDestroy(c->gtail);
  for (hipGraphExec_t g : c->gchunks) hipGraphExecDestroy(g);
  c->gexec = c->gtail = nullptr;
  c->gchunks.clear();
}

static hipEvent_t get_event(smdlt_ctx* c) {
  if (!c->pool.empty()) {
    hipEvent_t e = c->pool.back();
    c->pool.pop_back();
    return e;
  }
  hipEvent_t e;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

struct StageScope {
  smdlt_ctx* c;
  int st;
  hipEvent_t b = nullptr;
  StageScope(smdlt_ctx* c_, int s) : c(c_), st(s) {
    if (c->prof) {
      b = get_event(c);
      hipEventRecord(b, c->stream);
    }
  }
  ~StageScope() {
    if (c->prof && b) {
      hipEvent_t e = get_event(c);
      hipEventRecord(e, c->stream);
      c->timers.push_back({b, e, st, true});
    }
  }
};

static void collect_timers(smdlt_ctx* c) {
  if (c->timers.empty()) return;
  hipStreamSynchronize(c->stream);
  for (auto& t : c->timers) {
    float ms = 0;
    hipEventElapsedTime(&ms, t.a, t.b);
    if (t.stage >= 0) {
      c->stage_ms[t.stage] += ms;
      c->stage_cnt[t.stage] += t.calls;
    }
    if (t.own_a) c->pool.push_back(t.a);
    c->pool.push_back(t.b);
  }
  c->timers.clear();
}

static int pad_slots(int n) { return ((n + SLOT_ALIGN - 1) / SLOT_ALIGN) * SLOT_ALIGN; }

// Fill the lattice-dependent fields of a replica descriptor.
static void fill_lattice(smdlt_ctx* c, int r) {
  RepDev& R = c->rh[r];
  const HostLattice& h = c->lat[r];
  R.L = h.L;
  R.invL = 1.0 / h.L;
  R.L1xp = h.L1xp;
  R.L2xp = h.L2xp;
  R.L2yp = h.L2yp;
  R.delx = h.delx;
  R.ct = h.ct;
  R.st = h.st;
  R.ct0 = c->ct_prev[r];
  R.st0 = c->st_prev[r];
  double wm;
  reduce_basis(h.L1xp, 0.0, h.L2xp, h.L2yp, R.rb, R.rd, &wm);
  R.halfw = 0.5 * std::min(wm, h.L);
  R.hinv[0] = 1.0 / h.L1xp;
  R.hinv[1] = 1.0 / h.L2yp;
  R.hinv[2] = 1.0 / h.L;
  R.kb[0] = 2 * M_PI / h.L1xp;
  R.kb[1] = -2 * M_PI * h.L2xp / (h.L1xp * h.L2yp);
  R.kb[2] = 2 * M_PI / h.L2yp;
  R.kb[3] = 2 * M_PI / h.L;
  R.vol = fabs(h.L1xp * h.L2yp) * h.L;
  R.treal = h.treal;
  // cell grid for the neighbour build at this lattice
  double w[2];
  reduced_widths(h, w);
  double cut = std::max(sqrt(R.rlist2), R.rv + 1.0);
  int n1 = std::max(1, (int)floor(w[0] / cut)), n2 = std::max(1, (int)floor(w[1] / cut));
  int n3 = std::max(1, (int)floor(h.L / cut));
  while ((long long)n1 * n2 * n3 > c->max_w_cells[r]) {
    if (n1 >= n2 && n1 >= n3 && n1 > 1)
      n1--;
    else if (n2 >= n3 && n2 > 1)
      n2--;
    else
      n3--;
  }
  R.ncell[0] = n1;
  R.ncell[1] = n2;
  R.ncell[2] = n3;
}

static int upload_reps(smdlt_ctx* c) {
  HCK(hipMemcpyAsync(c->d_reps, c->rh.data(), sizeof(RepDev) * c->nrep, hipMemcpyHostToDevice, c->stream));
  return 0;
}

static void set_active(smdlt_ctx* c, int rb, int re) {
  for (int r = 0; r < c->nrep; ++r) c->rh[r].active = (r >= rb && r < re) ? 1 : 0;
  c->cur_rep_begin = rb;
  c->cur_rep_end = re;
}

template <typename T> static int alloc_dev(T** p, size_t n) {
  if (n == 0) n = 1;
  hipError_t e = hipMalloc((void**)p, n * sizeof(T));
  if (e != hipSuccess) {
    last_hip_error() = hipGetErrorString(e);
    return SMDLT_E_NOMEM;
  }
  hipMemset(*p, 0, n * sizeof(T));
  return 0;
}

extern "C" {

const char* smdlt_version(void) { return "smdlt-mi355x 0.1 (gfx950)"; }

const char* smdlt_strerror(int code) {
  switch (code) {
    case SMDLT_OK: return "ok";
    case SMDLT_E_ARG: return "bad argument";
    case SMDLT_E_HIP: return last_hip_error().empty() ? "HIP runtime error" : last_hip_error().c_str();
    case SMDLT_E_FENE_OVERSTRETCH: return "FENE overstretch";
    case SMDLT_E_LANCZOS_NOCONV: return "Lanczos did not converge";
    case SMDLT_E_NLIST_OVERFLOW: return "neighbour list capacity exceeded";
    case SMDLT_E_GEOMETRY: return "cutoff too large for the box";
    case SMDLT_E_STATE: return "call order violated";
    case SMDLT_E_NOMEM: return "device allocation failed";
    case SMDLT_E_MC_STREAM: return "MC uniform stream exhausted";
    case SMDLT_E_NONFINITE: return "non-finite displacement (not applied)";
    default: return "unknown error";
  }
}

void smdlt_default_params(smdlt_params* p) {
  memset(p, 0, sizeof(*p));
  p->nb = 50;
  p->nc = 20;
  p->L = 60.0;
  p->flow_type = SMDLT_FLOW_SHEAR;
  p->edot = 0.1;
  p->dt = 0.0005;
  p->spring_fene = 0;
  p->epsilon = 0.1;
  p->kappab = 0.0;
  p->kappas = 200.0;
  p->qmax = 2.0;
  p->lambda_b = 3.59;
  p->lambda_d = 5.0;
  p->barrier = 3.0;
  p->barrier2 = 3.0;
  p->lj_ext_form = 0;
  p->pbc_compat = 0;  // exact Lees-Edwards wrap; 1 reproduces the reference's +-delx shift (parity runs)
  p->precision = SMDLT_PREC_MIXED;
  p->noise_mode = SMDLT_NOISE_LANCZOS;
  p->mobility_tol = 1e-5;
  p->lanczos_tol = 1e-6;
  p->lanczos_maxit = 30;
  p->seed = 0x5EED0000ull;
  p->hop_radius = 3.0;
  p->feature_radius = 4.0;
  p->proposal_floor = 0.05;
}

int smdlt_num_beads(const smdlt_ctx* c, int replica) {
  if (!c) return SMDLT_E_ARG;
  if (replica < 0) {
    int s = 0;
    for (int r = 0; r < c->nrep; ++r) s += c->rh[r].n;
    return s;
  }
  if (replica >= c->nrep) return SMDLT_E_ARG;
  return c->rh[replica].n;
}

// Mobility settings.  Error model: the real-space tail |t1|+|t2| at rc is held
// at 0.1 tol, and the grid part is given a quarter of tol: that is its RMS over
// beads, and one bead's self term can be 2-3 times the RMS.  The grid error was
// measured as the relative norm error of the FP64 product at a coarse grid
// against a fine one (h alpha = 0.12, p = 8; same alpha and rc; random v;
// chain systems of 1000-8000 beads at 0.005-0.04 beads per unit volume,
// alpha 0.12-0.7): it reaches 5e-6 at h alpha = 0.32 (p=8), 0.235 (p=6),
// 0.095 (p=4) for alpha <= 0.4, grows as (h alpha)^(1.6p - 2.5), and at larger
// alpha (small boxes) the admissible h alpha shrinks about as alpha^-0.45.
// Among the (rc, p) candidates that fit the box, the cheapest by a cost model
// of the M.v kernels (pairs, grid passes, spline points) is chosen.
static double grid_h_alpha(int p, double tol, double alpha) {
  double c = (p == 4) ? 0.095 : (p == 6) ? 0.235 : 0.32;
  double s = pow(0.25 * tol / 5e-6, 1.0 / (1.6 * p - 2.5));
  if (alpha > 0.4) s *= pow(0.4 / alpha, 0.45);
  return c * s;
}

struct MobChoice {
  double rc, alpha;
  int K[3], order;
  double cost, nbr;
};

static MobChoice mob_candidate(double rc, int order, double tol, double skin, const double len[3], int n, double vol,
                               int nb, bool fp64) {
  MobChoice m;
  m.rc = rc;
  m.order = order;
  m.alpha = choose_alpha(rc, 0.1 * tol);
  double h = grid_h_alpha(order, tol, m.alpha) / m.alpha;
  for (int d = 0; d < 3; ++d) m.K[d] = next_fft_len((int)ceil(len[d] / h));
  double rl = rc + skin;
  double chain = std::min((double)nb, 2.0 * pow(rl / 2.0, 1.7));
  m.nbr = (double)n / vol * (4.0 / 3.0) * M_PI * rl * rl * rl + chain;
  double G = (double)m.K[0] * m.K[1] * m.K[2];
  // Costs (ns per M.v) fitted to MI355X kernel times of 128 replicas of the
  // uneven ensemble (0.68 M beads) over cutoffs 12-32 and orders 6 and 8: per
  // real-space list pair (within rc + skin) 0.0014, per grid point 0.047 (the
  // five FFT passes 0.039 and the gather's per-point part 0.0078), per stencil
  // point (bead x p^3) 0.0051 (the gather's per-bead part 0.0024 and the
  // interpolation 0.0027).  The list of a replica is capped at 2500 pairs per
  // bead, which bounds the bead-pair list memory.
  // FP64 contexts evaluate every pair's coefficients in closed form (erfc, exp)
  // on a bead-pair list; their model prices pairs as in the FP64 kernels.
  if (fp64)
    m.cost = 0.02 * n * m.nbr + 0.15 * G + 0.01 * n * order * order * order;
  else
    m.cost = 0.0014 * n * m.nbr + 0.047 * G + 0.0051 * n * order * order * order;
  bool bad = false;
  for (int d = 0; d < 3; ++d) bad = bad || m.K[d] > KMAX;
  if (bad || m.nbr > 2500) m.cost = 1e300;
  return m;
}

static double rtab_knot_up(double x);

static int choose_mobility(smdlt_ctx* c, int r) {
  const smdlt_params& p = c->P[r];
  RepDev& R = c->rh[r];
  HostLattice h = c->lat[r];
  // geometry over one lattice period: smallest reduced width, longest lattice vectors
  double wmin = 1e300, a1max = 0, a2max = 0;
  int nsamp = (h.tp == INT_MAX) ? 1 : 65;
  for (int k = 0; k < nsamp; ++k) {
    uint64_t t = (h.tp == INT_MAX) ? 0 : (uint64_t)((double)k / (nsamp - 1) * (h.tp - 1));
    h.update(t);
    double w[2];
    reduced_widths(h, w);
    wmin = std::min(wmin, std::min(std::min(w[0], w[1]), h.L));
    a1max = std::max(a1max, fabs(h.L1xp));
    a2max = std::max(a2max, sqrt(h.L2xp * h.L2xp + h.L2yp * h.L2yp));
  }
  double tol = p.mobility_tol > 0 ? p.mobility_tol : 1e-5;
  double skin = p.skin > 0 ? p.skin : 1.0;
  double rcmax = 0.5 * wmin - skin - 1e-6;
  double len[3] = {a1max, a2max, h.L};
  int n = R.n;
  double vol = p.L * p.L * p.L;
  MobChoice best;
  best.cost = 1e301;
  const double rcs[] = {5, 6, 7, 8, 10, 12, 14, 16, 20, 24, 28, 32, 36, 40, 44, 48};
  const int orders[] = {4, 6, 8};
  for (double rc : rcs) {
    if (p.rc_real > 0) rc = p.rc_real;
    if (rc > rcmax) continue;
    if (c->prec == SMDLT_PREC_FP64 && rc > 32 && p.rc_real <= 0) continue;  // the FP64 model's fitted range
    for (int o : orders) {
      if (p.spline_order > 0 && o != p.spline_order) continue;
      MobChoice m = mob_candidate(rc, o, tol, skin, len, n, vol, p.nb, c->prec == SMDLT_PREC_FP64);
      if (m.cost < best.cost) best = m;
    }
    if (p.rc_real > 0) break;
  }
  if (best.cost >= 1e300) {
    if (p.rc_real > 0 && p.rc_real > rcmax) return SMDLT_E_GEOMETRY;
    if (rcmax < 2.0) return SMDLT_E_GEOMETRY;
    best = mob_candidate(std::min(rcmax, 8.0), p.spline_order > 0 ? p.spline_order : 8, tol, skin, len, n, vol, p.nb,
                         c->prec == SMDLT_PREC_FP64);
  }
  if (p.ewald_alpha > 0) {
    best.alpha = p.ewald_alpha;
    double hh = grid_h_alpha(best.order, tol, best.alpha) / best.alpha;
    for (int d = 0; d < 3; ++d) best.K[d] = next_fft_len((int)ceil(len[d] / hh));
  }
  if (p.grid[0] > 0 && p.grid[1] > 0 && p.grid[2] > 0)
    for (int d = 0; d < 3; ++d) best.K[d] = p.grid[d];
  for (int d = 0; d < 3; ++d)
    if (!good_fft_len(best.K[d])) return SMDLT_E_ARG;
  if (best.order != 4 && best.order != 6 && best.order != 8) return SMDLT_E_ARG;
  double rc = best.rc;
  R.K[0] = best.K[0];
  R.K[1] = best.K[1];
  R.K[2] = best.K[2];
  R.order = best.order;
  R.alpha = best.alpha;
  R.rc2 = c->prec == SMDLT_PREC_FP64 ? rc * rc : rtab_knot_up(rc * rc);
  R.rlist2 = (sqrt(R.rc2) + skin) * (sqrt(R.rc2) + skin);
  R.skin = skin;
  R.selfconst = ewald_selfconst(best.alpha);
  c->nbr_est = std::max(c->nbr_est, best.nbr);
  // cell capacity over the period
  double cut = std::max(rc + skin, 6.0 + 1.0);
  double wmax = 0;
  h = c->lat[r];
  for (int k = 0; k < nsamp; ++k) {
    uint64_t t = (h.tp == INT_MAX) ? 0 : (uint64_t)((double)k / (nsamp - 1) * (h.tp - 1));
    h.update(t);
    double w[2];
    reduced_widths(h, w);
    wmax = std::max(wmax, std::max(w[0], w[1]));
  }
  int nmaxc = std::max(1, (int)floor(wmax / cut));
  int nz = std::max(1, (int)floor(h.L / cut));
  c->max_w_cells[r] = nmaxc * nmaxc * nz;
  return 0;
}

// Beenakker real-space coefficients for r >= 2 (a = 1): M = t1 I + t2 d d^T / r^2
// (S/main26.c:1727-1745), FP64 on the host for the mixed-precision tables.
static void rpy_real_host(double r, double alpha, double& t1, double& t2) {
  const double rr = r * r, ir = 1.0 / r, ir3 = ir / rr;
  const double a2 = alpha * alpha, a3 = a2 * alpha, a5 = a3 * a2, a7 = a5 * a2;
  const double ec = erfc(alpha * r), g = exp(-a2 * rr) / sqrt(M_PI);
  const double C1 = 0.75 * ir + 0.5 * ir3, C3 = 0.75 * ir - 1.5 * ir3;
  const double C2 = 4 * a7 * rr * rr + 3 * a3 * rr - 20 * a5 * rr - 4.5 * alpha + 14 * a3 + alpha / rr;
  const double C4 = -4 * a7 * rr * rr - 3 * a3 * rr + 16 * a5 * rr + 1.5 * alpha - 2 * a3 - 3 * alpha / rr;
  t1 = C1 * ec + C2 * g;
  t2 = C3 * ec + C4 * g;
}

// Cubic coefficients per table segment (rtab_locate: r^2 from the float bits
// (RTAB_BASE + i) << RTAB_SHIFT to those of i + 1, position s in [1, 2) linear in
// r^2) of t1(r) and t2(r)/r^2, interpolating at the 4 Chebyshev nodes of the
// segment (FP64 solve).  Below r = 2 the exact RPY overlap form replaces the far
// form, written with erf so the 1/r^3 terms cancel analytically; r^2 = 4 is an
// octave start and so a segment knot.
static double rtab_knot(int i) {
  const uint32_t b = (uint32_t)(RTAB_BASE + (unsigned)i) << RTAB_SHIFT;
  float f;
  memcpy(&f, &b, 4);
  return (double)f;
}

// Segments covering r^2 in [1, rc2), plus the zero entry.  Mixed-precision
// cut-offs are table knots (rtab_knot_up), so the zero entry starts at rc2 and
// the kernels' clamp of r^2 to [1, rc2] is their cut-off test.
static int rtab_segments(double rc2) {
  const float rcf = (float)rc2;
  uint32_t b;
  memcpy(&b, &rcf, 4);
  const int idx = (int)((b >> RTAB_SHIFT) - RTAB_BASE);  // the segment holding rc2
  const bool knot = (b & ((1u << RTAB_SHIFT) - 1u)) == 0u;
  return std::max(1, knot ? idx : idx + 1) + 1;
}

// The smallest table knot >= x (x >= 1): r^2 cut-offs of the mixed-precision
// kernels sit on knots, raising rc by less than 1/64 of itself.
static double rtab_knot_up(double x) {
  float f = (float)x;
  if ((double)f < x) f = nextafterf(f, INFINITY);
  uint32_t b;
  memcpy(&b, &f, 4);
  if (b & ((1u << RTAB_SHIFT) - 1u)) b = ((b >> RTAB_SHIFT) + 1u) << RTAB_SHIFT;
  memcpy(&f, &b, 4);
  return (double)f;
}

static void build_rtab(double alpha, int ntab, float* out) {
  double tn[4], V[4][4];
  for (int k = 0; k < 4; ++k) {
    tn[k] = 0.5 * (1.0 - cos((2 * k + 1) * M_PI / 8.0));
    for (int p = 0; p < 4; ++p) V[k][p] = pow(1.0 + tn[k], p);
  }
  // invert the 4x4 Vandermonde once (Gauss-Jordan)
  double A[4][8];
  for (int k = 0; k < 4; ++k)
    for (int p = 0; p < 8; ++p) A[k][p] = p < 4 ? V[k][p] : (p - 4 == k ? 1.0 : 0.0);
  for (int col = 0; col < 4; ++col) {
    int piv = col;
    for (int k = col + 1; k < 4; ++k)
      if (fabs(A[k][col]) > fabs(A[piv][col])) piv = k;
    for (int p = 0; p < 8; ++p) std::swap(A[col][p], A[piv][p]);
    double d = A[col][col];
    for (int p = 0; p < 8; ++p) A[col][p] /= d;
    for (int k = 0; k < 4; ++k)
      if (k != col) {
        double f = A[k][col];
        for (int p = 0; p < 8; ++p) A[k][p] -= f * A[col][p];
      }
  }
  for (int p = 0; p < 8; ++p) out[8 * (ntab - 1) + p] = 0.f;
  for (int i = 0; i < ntab - 1; ++i) {
    double f1[4], f2[4];
    for (int k = 0; k < 4; ++k) {
      const double lo = rtab_knot(i), hi = rtab_knot(i + 1);
      double r = sqrt(lo + (hi - lo) * tn[k]), t1, t2;
      if (r >= 2.0) {
        rpy_real_host(r, alpha, t1, t2);
      } else {
        const double rr = r * r, ir = 1.0 / r, ir3 = ir / rr;
        const double a2 = alpha * alpha, a3 = a2 * alpha, a5 = a3 * a2, a7 = a5 * a2;
        const double ef = erf(alpha * r), g = exp(-a2 * rr) / sqrt(M_PI);
        const double C1 = 0.75 * ir + 0.5 * ir3, C3 = 0.75 * ir - 1.5 * ir3;
        const double C2 = 4 * a7 * rr * rr + 3 * a3 * rr - 20 * a5 * rr - 4.5 * alpha + 14 * a3 + alpha / rr;
        const double C4 = -4 * a7 * rr * rr - 3 * a3 * rr + 16 * a5 * rr + 1.5 * alpha - 2 * a3 - 3 * alpha / rr;
        t1 = (1.0 - 0.28125 * r) - C1 * ef + C2 * g;
        t2 = 0.09375 * r - C3 * ef + C4 * g;
      }
      f1[k] = t1;
      f2[k] = t2 / (r * r);
    }
    for (int p = 0; p < 4; ++p) {
      double c1 = 0, c2 = 0;
      for (int k = 0; k < 4; ++k) {
        c1 += A[p][4 + k] * f1[k];
        c2 += A[p][4 + k] * f2[k];
      }
      out[8 * i + p] = (float)c1;
      out[8 * i + 4 + p] = (float)c2;
    }
  }
}

int smdlt_create(const smdlt_params* params, int n_replicas, int device, smdlt_ctx** out) {
  if (!params || n_replicas <= 0 || !out) return SMDLT_E_ARG;
  *out = nullptr;
  HCK(hipSetDevice(device));
  {
    int st0 = resolve_scratch(device);  // smdlt_mc_sweep_noalloc's buffer on this device
    if (st0) return st0;
  }
  smdlt_ctx* c = new smdlt_ctx();
  c->device = device;
  c->nrep = n_replicas;
  c->P.assign(params, params + n_replicas);
  c->prec = params[0].precision;
  c->lat.resize(n_replicas);
  c->rh.assign(n_replicas, RepDev());
  memset(c->rh.data(), 0, sizeof(RepDev) * n_replicas);
  c->ct_prev.assign(n_replicas, 1.0);
  c->st_prev.assign(n_replicas, 0.0);
  c->max_w_cells.assign(n_replicas, 1);
  c->nq.assign(n_replicas, 0);
  c->q0.assign(n_replicas, 0);
  c->last_iters.assign(n_replicas, 0);
  int slot = 0, wave = 0, maxit = 1;
  long long g = 0;
  int bsp_off = 0, cells = 0;
  for (int r = 0; r < n_replicas; ++r) {
    const smdlt_params& p = params[r];
    if (p.nb <= 0 || p.nc <= 0 || p.L <= 0 || p.dt <= 0 || p.precision != c->prec) {
      delete c;
      return SMDLT_E_ARG;
    }
    RepDev& R = c->rh[r];
    R.n = p.nb * p.nc;
    R.nb = p.nb;
    R.nc = p.nc;
    R.slot0 = slot;
    R.wave0 = wave;
    int ps = pad_slots(R.n);
    R.nwaves = ps / SLOT_ALIGN;
    slot += ps;
    wave += R.nwaves;
    R.flow = p.flow_type;
    R.pbc_compat = p.pbc_compat;
    R.lj_ext = p.lj_ext_form;
    R.spring_fene = p.spring_fene;
    R.dt = p.dt;
    R.edot = (p.flow_type == 0) ? 0.0 : p.edot;
    R.kappas = p.kappas;
    R.qmax = p.qmax;
    R.kappab = p.kappab;
    R.epsilon = p.epsilon;
    R.lambda_b = p.lambda_b;
    R.kappa_d = 1.0 / p.lambda_d;
    R.barrier = p.barrier;
    R.barrier2 = p.barrier2;
    {
      double hop = p.hop_radius > 0 ? p.hop_radius : 3.0, fr = p.feature_radius > 0 ? p.feature_radius : 4.0;
      R.hop2 = hop * hop;
      R.feat_r = fr;
      R.nbhd2 = (2 * hop + fr + 0.5) * (2 * hop + fr + 0.5);  // every bead any reverse feature can see
      R.pfloor = (p.proposal_floor > 0 && p.proposal_floor <= 1) ? p.proposal_floor : 0.05;
    }
    R.r2cut = p.lj_ext_form ? 5.03968 : 25.0;
    R.rv = 6.0;
    R.rnew = 0.5 * (R.rv - sqrt(R.r2cut));
    R.sqrt2dt = sqrt(2.0 * p.dt);
    R.active = 1;
    c->lat[r].init(p.flow_type, p.L, R.edot, p.dt, p.pbc_compat);
    c->ct_prev[r] = c->lat[r].ct;
    c->st_prev[r] = c->lat[r].st;
    if (r > 0 && c->P[r].spline_order == 0) c->P[r].spline_order = c->rh[0].order;  // one order per context
    int st = choose_mobility(c, r);
    if (st) {
      delete c;
      return st;
    }
    R.grid0 = g;
    g += 3LL * R.K[0] * R.K[1] * R.K[2];
    R.bsp0 = bsp_off;
    bsp_off += R.K[0] + R.K[1] + R.K[2];
    R.cell0 = cells;
    cells += c->max_w_cells[r];
    maxit = std::max(maxit, p.lanczos_maxit > 0 ? p.lanczos_maxit : 30);
    fill_lattice(c, r);
  }
  if (maxit > LZ_MAXIT) {
    delete c;
    return SMDLT_E_ARG;
  }
  c->maxit = maxit;
  c->cap_rpy = std::max(64, (int)(1.5 * c->nbr_est) + 32);
  c->nst = slot;
  c->nwaves = wave;
  c->grid_total = g;
  c->ncells = cells;
  int st = 0;
  HCK(hipStreamCreate(&c->stream));
  size_t n3 = 3 * (size_t)c->nst;
  st |= alloc_dev(&c->d_reps, n_replicas);
  st |= alloc_dev(&c->d_slot_rep, c->nst);
  st |= alloc_dev(&c->d_seeds, n_replicas);
  st |= alloc_dev(&c->rx, c->nst);
  st |= alloc_dev(&c->ry, c->nst);
  st |= alloc_dev(&c->rz, c->nst);
  st |= alloc_dev(&c->dxv, c->nst);
  st |= alloc_dev(&c->dyv, c->nst);
  st |= alloc_dev(&c->dzv, c->nst);
  st |= alloc_dev(&c->dsp, n3);
  st |= alloc_dev(&c->f, n3);
  st |= alloc_dev(&c->z, n3);
  st |= alloc_dev(&c->mf, n3);
  st |= alloc_dev(&c->bz, n3);
  st |= alloc_dev(&c->tmp, n3);
  st |= alloc_dev(&c->tmp2, n3);
  st |= alloc_dev(&c->nrpy, c->nst);
  st |= alloc_dev(&c->nlj, c->nst);
  st |= alloc_dev(&c->lrpy, (size_t)c->cap_rpy * c->nst);
  st |= alloc_dev(&c->llj, (size_t)c->cap_lj * c->nst);
  st |= alloc_dev(&c->overflow, 1);
  st |= alloc_dev(&c->d_err, 4);
  st |= alloc_dev(&c->cell_of, c->nst);
  st |= alloc_dev(&c->cell_cnt, c->ncells + 1);
  st |= alloc_dev(&c->cell_start, c->ncells + 1);
  st |= alloc_dev(&c->cell_fill, c->ncells + 1);
  st |= alloc_dev(&c->cell_beads, c->nst);
  st |= alloc_dev(&c->cell_tmp, c->nst);
  st |= alloc_dev(&c->rep_flags, n_replicas);
  st |= alloc_dev(&c->out2, 2 * n_replicas);
  st |= alloc_dev(&c->aff, n_replicas);
  size_t esz = (c->prec == SMDLT_PREC_FP64) ? 8 : 4;
  {
    hipError_t e = hipMalloc(&c->grid, std::max<size_t>(1, (size_t)c->grid_total * esz));
    if (e != hipSuccess) st |= SMDLT_E_NOMEM;
    e = hipMalloc(&c->bsp, std::max<size_t>(1, (size_t)bsp_off * esz));
    if (e != hipSuccess) st |= SMDLT_E_NOMEM;
  }
  st |= alloc_dev(&c->V, (size_t)(c->maxit + 1) * n3);
  st |= alloc_dev(&c->W, n3);
  st |= alloc_dev(&c->partial, c->nwaves + 1);
  st |= alloc_dev(&c->partial2, c->nwaves + 1);
  st |= alloc_dev(&c->scal, (size_t)SCAL_STRIDE * n_replicas);
  st |= alloc_dev(&c->bdiv, n_replicas);
  st |= alloc_dev(&c->done, n_replicas);
  st |= alloc_dev(&c->iters, n_replicas);
  st |= alloc_dev(&c->mc_err, 4);
  st |= alloc_dev(&c->saved, 9 * (size_t)c->nst + 3 * (size_t)n_replicas);
  c->chain0.assign(n_replicas + 1, 0);
  for (int r = 0; r < n_replicas; ++r) {
    const int ncr = c->rh[r].n / c->rh[r].nb;
    c->chain0[r + 1] = c->chain0[r] + ncr;
    c->max_chains = std::max(c->max_chains, ncr);
  }
  c->nchains = c->chain0[n_replicas];
  st |= alloc_dev(&c->vir, 3 * (size_t)c->nst);
  st |= alloc_dev(&c->qrel, 3 * (size_t)c->nst);
  st |= alloc_dev(&c->tau, 3 * (size_t)n_replicas);
  st |= alloc_dev(&c->d_chain0, n_replicas + 1);
  st |= alloc_dev(&c->obs, 10 * (size_t)std::max(1, c->nchains));
  if (hipHostMalloc((void**)&c->h_flags, 8 * (size_t)n_replicas + 4, hipHostMallocDefault) != hipSuccess) {
    c->h_flags = nullptr;
    st |= SMDLT_E_NOMEM;
  }
  if (st) {
    smdlt_destroy(c);
    return SMDLT_E_NOMEM;
  }
  // Lanczos scalar blocks start at zero (no previous solve: first square root at iteration 2)
  HCK(hipMemset(c->scal, 0, sizeof(double) * SCAL_STRIDE * (size_t)n_replicas));
  HCK(hipMemset(c->iters, 0, sizeof(int) * (size_t)n_replicas));
  HCK(hipMemset(c->tau, 0, sizeof(double) * 3 * (size_t)n_replicas));
  HCK(hipMemcpy(c->d_chain0, c->chain0.data(), sizeof(int) * (n_replicas + 1), hipMemcpyHostToDevice));
  // slot -> replica map, seeds
  std::vector<int> sr(c->nst, -1);
  std::vector<uint64_t> seeds(n_replicas);
  for (int r = 0; r < n_replicas; ++r) {
    for (int s = 0; s < pad_slots(c->rh[r].n); ++s) sr[c->rh[r].slot0 + s] = r;
    seeds[r] = params[r].seed;
  }
  HCK(hipMemcpy(c->d_slot_rep, sr.data(), sizeof(int) * c->nst, hipMemcpyHostToDevice));
  HCK(hipMemcpy(c->d_seeds, seeds.data(), sizeof(uint64_t) * n_replicas, hipMemcpyHostToDevice));
  // twiddles exp(-2 pi i t / K), one table per distinct line length, and B-spline moduli
  {
    std::vector<int> lens;
    std::vector<double> twd;
    for (int r = 0; r < n_replicas; ++r)
      for (int d = 0; d < 3; ++d) {
        const int K = c->rh[r].K[d];
        int off = -1, acc = 0;
        for (int L : lens) {
          if (L == K) off = acc;
          acc += L;
        }
        if (off < 0) {
          off = acc;
          lens.push_back(K);
          for (int t = 0; t < K; ++t) {
            twd.push_back(cos(2 * M_PI * t / K));
            twd.push_back(-sin(2 * M_PI * t / K));
          }
        }
        c->rh[r].twoff[d] = off;
      }
    const size_t ntw = twd.size() / 2;
    if (hipMalloc(&c->tw, ntw * 2 * esz) != hipSuccess) {
      smdlt_destroy(c);
      return SMDLT_E_NOMEM;
    }
    std::vector<double> bs(bsp_off);
    for (int r = 0; r < n_replicas; ++r) {
      int off = c->rh[r].bsp0;
      for (int d = 0; d < 3; ++d) {
        std::vector<double> m;
        bspline_moduli(c->rh[r].K[d], c->rh[r].order, m);
        std::copy(m.begin(), m.end(), bs.begin() + off);
        off += c->rh[r].K[d];
      }
    }
    if (esz == 8) {
      HCK(hipMemcpy(c->tw, twd.data(), 16 * ntw, hipMemcpyHostToDevice));
      HCK(hipMemcpy(c->bsp, bs.data(), 8 * bs.size(), hipMemcpyHostToDevice));
    } else {
      std::vector<float> tf(twd.begin(), twd.end()), bf(bs.begin(), bs.end());
      HCK(hipMemcpy(c->tw, tf.data(), 8 * ntw, hipMemcpyHostToDevice));
      HCK(hipMemcpy(c->bsp, bf.data(), 4 * bf.size(), hipMemcpyHostToDevice));
    }
  }
  // FFT task tables, grouped by replica
  {
    std::vector<FftTask> vx, vy, vz, vxy;
    // The fused x-y pass saves two launches and a grid round trip, which pays on
    // small grids; on large batches the separate passes keep more waves per CU
    // and were measured ~5% faster, so fusing stops above XY_MAX_POINTS in total.
    double gtot = 0;
    for (int r = 0; r < n_replicas; ++r) gtot += (double)c->rh[r].K[0] * c->rh[r].K[1] * c->rh[r].K[2];
    c->tx0.assign(n_replicas + 1, 0);
    c->ty0.assign(n_replicas + 1, 0);
    c->tz0.assign(n_replicas + 1, 0);
    c->txy0.assign(n_replicas + 1, 0);
    // The z pass reads runs of nlz consecutive lines, so longer runs coalesce
    // better, but on small grids a large tile leaves too few workgroups: the
    // 40x48x40 grid has 39 with 1024 elements (7% slower per step than 512)
    // and 160 with 256 (3% faster than 512).
    auto z_tasks = [&](int tile) {
      long long n = 0;
      for (int r = 0; r < n_replicas; ++r) {
        const int nl = std::max(1, std::min(FFT_MAXL, tile / c->rh[r].K[2]));
        n += ((long long)c->rh[r].K[1] * (c->rh[r].K[0] / 2) + nl - 1) / nl;
      }
      return n;
    };
    // A z task holds nl >= 1 whole lines of every replica, so a tile is only
    // eligible when the longest z line fits it (FFT_TILE_Z >= KMAX always does):
    // a shorter tile would overrun the LDS lines and the butterfly slots.
    const bool fp64 = c->prec == SMDLT_PREC_FP64;
    int kzmax = 0;
    for (int r = 0; r < n_replicas; ++r) kzmax = std::max(kzmax, c->rh[r].K[2]);
    c->ztile = (!fp64 && z_tasks(FFT_TILE_Z_LARGE) >= FFT_Z_LARGE_MIN_TASKS) ? FFT_TILE_Z_LARGE
               : (z_tasks(FFT_TILE_Z) >= FFT_Z_MIN_TASKS || kzmax > FFT_TILE_Z_SMALL) ? FFT_TILE_Z
                                                                                      : FFT_TILE_Z_SMALL;
    static_assert(FFT_TILE_Z >= KMAX && FFT_TILE_X >= KMAX && FFT_TILE_X64 >= KMAX && FFT_TILE_Y64 >= KMAX &&
                      FFT_TILE_Y >= KMAX,
                  "every FFT tile except FFT_TILE_Z_SMALL must hold a whole line of the longest length");
    const int ztile = c->ztile, ytile = fp64 ? FFT_TILE_Y64 : FFT_TILE_Y, xtile = fp64 ? FFT_TILE_X64 : FFT_TILE_X;
    for (int r = 0; r < n_replicas; ++r) {
      const RepDev& R = c->rh[r];
      int K0 = R.K[0], K1 = R.K[1], K2 = R.K[2], Kh = K0 / 2;
      c->tx0[r] = (int)vx.size();
      c->ty0[r] = (int)vy.size();
      c->tz0[r] = (int)vz.size();
      c->txy0[r] = (int)vxy.size();
      int nlx = std::max(1, std::min(FFT_MAXL, xtile / K0));
      int nly = std::max(1, std::min(FFT_MAXL, ytile / K1));
      int nlz = std::max(1, std::min(FFT_MAXL, ztile / K2));
      const bool fused = K0 * K1 / 2 <= XY_ELEMS && gtot <= XY_MAX_POINTS;  // the plane fits one workgroup's LDS
      for (int comp = 0; comp < 3; ++comp) {
        if (fused) {
          for (int zz = 0; zz < K2; ++zz) vxy.push_back({r, comp, zz, 1});
          continue;
        }
        int pairs = K2 * K1 / 2;
        for (int l0 = 0; l0 < pairs; l0 += nlx) vx.push_back({r, comp, l0, std::min(nlx, pairs - l0)});
        int lines = K2 * Kh;
        for (int l0 = 0; l0 < lines; l0 += nly) vy.push_back({r, comp, l0, std::min(nly, lines - l0)});
      }
      int zl = K1 * Kh;
      for (int l0 = 0; l0 < zl; l0 += nlz) vz.push_back({r, -1, l0, std::min(nlz, zl - l0)});
    }
    // FFT chunks: consecutive replicas with at most FFT_CHUNK_BYTES of grid (a
    // larger replica is a chunk of its own)
    c->fch_rep.clear();
    double acc = 0;
    for (int r = 0; r < n_replicas; ++r) {
      const double gb = (fp64 ? 24.0 : 12.0) * c->rh[r].K[0] * c->rh[r].K[1] * c->rh[r].K[2];
      if (acc > 0 && acc + gb > FFT_CHUNK_BYTES) {
        c->fch_rep.push_back(r);
        acc = 0;
      }
      acc += gb;
    }
    c->tx0[n_replicas] = (int)vx.size();
    c->ty0[n_replicas] = (int)vy.size();
    c->tz0[n_replicas] = (int)vz.size();
    c->txy0[n_replicas] = (int)vxy.size();
    c->ntx = (int)vx.size();
    c->nty = (int)vy.size();
    c->ntz = (int)vz.size();
    c->ntxy = (int)vxy.size();
    st |= alloc_dev(&c->tx, std::max<size_t>(1, vx.size()));
    st |= alloc_dev(&c->ty, std::max<size_t>(1, vy.size()));
    st |= alloc_dev(&c->tz, vz.size());
    st |= alloc_dev(&c->txy, std::max<size_t>(1, vxy.size()));
    if (st) {
      smdlt_destroy(c);
      return SMDLT_E_NOMEM;
    }
    if (!vx.empty()) HCK(hipMemcpy(c->tx, vx.data(), sizeof(FftTask) * vx.size(), hipMemcpyHostToDevice));
    if (!vy.empty()) HCK(hipMemcpy(c->ty, vy.data(), sizeof(FftTask) * vy.size(), hipMemcpyHostToDevice));
    HCK(hipMemcpy(c->tz, vz.data(), sizeof(FftTask) * vz.size(), hipMemcpyHostToDevice));
    if (!vxy.empty()) HCK(hipMemcpy(c->txy, vxy.data(), sizeof(FftTask) * vxy.size(), hipMemcpyHostToDevice));
  }
  // spreading bricks (8x8x8 grid points), grouped by replica
  {
    std::vector<BrickTask> vb;
    c->bk0.assign(n_replicas + 1, 0);
    for (int r = 0; r < n_replicas; ++r) {
      RepDev& R = c->rh[r];
      c->bk0[r] = (int)vb.size();
      R.brick0 = (int)vb.size();
      for (int d = 0; d < 3; ++d) R.nbrick[d] = R.K[d] / BRICK;
      for (int b2 = 0; b2 < R.nbrick[2]; ++b2)
        for (int b1 = 0; b1 < R.nbrick[1]; ++b1)
          for (int b0 = 0; b0 < R.nbrick[0]; ++b0) vb.push_back({r, b0, b1, b2});
    }
    c->bk0[n_replicas] = (int)vb.size();
    c->nbricks = (int)vb.size();
    st |= alloc_dev(&c->bricks, vb.size());
    st |= alloc_dev(&c->brick_cnt, vb.size() + 1);
    st |= alloc_dev(&c->brick_start, vb.size() + 1);
    st |= alloc_dev(&c->brick_fill, vb.size() + 1);
    st |= alloc_dev(&c->brick_beads, c->nst);
    st |= alloc_dev(&c->bbase, 3 * (size_t)c->nst);
    if (hipMalloc(&c->bw, (size_t)BW_STRIDE * c->nst * esz) != hipSuccess) st |= SMDLT_E_NOMEM;
    // zeroed: the cluster pass loads (and discards) the slots that follow a
    // cluster, which must hold finite values
    if (esz == 4) st |= alloc_dev(&c->p32, c->nst) | alloc_dev(&c->v32, c->nst);
    if (esz == 4)  // every bead is a candidate of 8 bricks
      st |= alloc_dev(&c->bcand_cnt, vb.size() + 1) | alloc_dev(&c->bcand_start, vb.size() + 1) |
            alloc_dev(&c->bcand, 8 * (size_t)c->nst);
    if (st) {
      smdlt_destroy(c);
      return SMDLT_E_NOMEM;
    }
    if (esz == 4) {  // real-space coefficient tables, one per replica
      int tot = 0;
      for (int r = 0; r < n_replicas; ++r) {
        RepDev& R = c->rh[r];
        R.ntab = rtab_segments(R.rc2);
        R.tab0 = tot;
        tot += 2 * R.ntab;
      }
      std::vector<float> tab(4 * (size_t)tot);
      for (int r = 0; r < n_replicas; ++r) build_rtab(c->rh[r].alpha, c->rh[r].ntab, tab.data() + 4 * c->rh[r].tab0);
      if (hipMalloc(&c->rtab, 16 * (size_t)tot) != hipSuccess) {
        smdlt_destroy(c);
        return SMDLT_E_NOMEM;
      }
      HCK(hipMemcpy(c->rtab, tab.data(), 16 * (size_t)tot, hipMemcpyHostToDevice));
      // clusters: runs of <= 8 consecutive beads of one chain
      std::vector<int> cf, cn, scl(c->nst, -1);
      int maxncl = 0;
      for (int r = 0; r < n_replicas; ++r) {
        RepDev& R = c->rh[r];
        R.cl0 = (int)cf.size();
        for (int ch = 0; ch < R.nc; ++ch)
          for (int j0 = 0; j0 < R.nb; j0 += 8) {
            int f = R.slot0 + ch * R.nb + j0, k = std::min(8, R.nb - j0);
            for (int q = 0; q < k; ++q) scl[f + q] = (int)cf.size();
            cf.push_back(f);
            cn.push_back(k);
          }
        R.ncl = (int)cf.size() - R.cl0;
        maxncl = std::max(maxncl, R.ncl);
      }
      if (c->nst > (1 << 27) - 8) {  // cluster-list entries pack the first slot in 27 bits
        smdlt_destroy(c);
        return SMDLT_E_ARG;
      }
      c->ncl_total = (int)cf.size();
      c->cl_words = (maxncl + 31) / 32;
      c->capc = std::min(maxncl, c->cap_rpy / 2 + 64);
      if (4LL * c->cl_words > 65536 || alloc_dev(&c->cl_first, c->ncl_total) || alloc_dev(&c->cl_cnt, c->ncl_total) ||
          alloc_dev(&c->slot_cl, c->nst) || alloc_dev(&c->nclp, c->ncl_total) ||
          alloc_dev(&c->adot_cl, c->ncl_total) ||
          alloc_dev(&c->lcl, (size_t)c->capc * c->ncl_total + CL_LIST_PAD)) {
        smdlt_destroy(c);
        return SMDLT_E_NOMEM;
      }
      HCK(hipMemcpy(c->cl_first, cf.data(), 4 * cf.size(), hipMemcpyHostToDevice));
      HCK(hipMemcpy(c->cl_cnt, cn.data(), 4 * cn.size(), hipMemcpyHostToDevice));
      HCK(hipMemcpy(c->slot_cl, scl.data(), 4 * scl.size(), hipMemcpyHostToDevice));
    }
    HCK(hipMemcpy(c->bricks, vb.data(), sizeof(BrickTask) * vb.size(), hipMemcpyHostToDevice));
  }
  set_active(c, 0, n_replicas);
  if (upload_reps(c)) {
    smdlt_destroy(c);
    return SMDLT_E_HIP;
  }
  HCK(hipStreamSynchronize(c->stream));
  *out = c;
  return 0;
}

void smdlt_destroy(smdlt_ctx* c) {
  if (!c) return;
  hipSetDevice(c->device);
  if (c->stream) hipStreamSynchronize(c->stream);
  destroy_graphs(c);
  if (c->h_flags) hipHostFree(c->h_flags);
  void* ptrs[] = {c->adot_cl, c->tea_beta, c->tea_C, c->tea_diag, c->cl_first, c->cl_cnt, c->slot_cl, c->lcl, c->nclp, c->rtab, c->p32, c->v32, c->bcand_cnt, c->bcand_start, c->bcand, c->mh_feats, c->mh_log, c->mlp_ws, c->mlp_x, c->mlp_y, c->mlp_part, c->mlp_out, c->saved, c->vir, c->tau, c->obs, c->d_chain0, c->qdisp, c->qdisp_saved, c->qrel, c->d_reps, c->d_slot_rep, c->d_seeds, c->rx, c->ry, c->rz, c->dxv, c->dyv, c->dzv, c->dsp, c->f,
                  c->z, c->mf, c->bz, c->tmp, c->tmp2, c->nrpy, c->lrpy, c->nlj, c->llj, c->overflow, c->d_err,
                  c->cell_of, c->cell_cnt, c->cell_start, c->cell_fill, c->cell_beads, c->cell_tmp, c->rep_flags, c->out2, c->aff,
                  c->partial2, c->grid, c->bsp, c->tw, c->tx, c->ty, c->tz, c->txy, c->bricks, c->bbase, c->brick_cnt, c->brick_start,
                  c->brick_fill, c->brick_beads, c->bw, c->V, c->W, c->partial, c->scal, c->bdiv, c->done,
                  c->iters, c->qtrack, c->qsign, c->qold, c->qidx, c->mc_err, c->mlp_w, c->occ, c->log_off,
                  c->mc_start, c->mc_jf, c->mc_de, c->mc_state, c->mh_cnt, c->mh_off, c->mh_list, c->mh_V,
                  c->mh_hcnt, c->mh_hoff, c->mh_hlist,
                  c->mh_moves, c->mh_nmoves, c->mh_done, c->mh_stop};
  for (void* p : ptrs)
    if (p) hipFree(p);
  for (auto& t : c->timers) {
    hipEventDestroy(t.a);
    hipEventDestroy(t.b);
  }
  for (auto e : c->pool) hipEventDestroy(e);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
}

int smdlt_get_mobility_config(const smdlt_ctx* c, int r, double out[9]) {
  if (!c || r < 0 || r >= c->nrep || !out) return SMDLT_E_ARG;
  const RepDev& R = c->rh[r];
  out[7] = 0.0;  // real-space list entries of the replica (after the last build)
  out[8] = 0.0;  // cluster-pair list entries (mixed precision)
  if (c->lists_built && R.n > 0) {
    HCK(hipSetDevice(c->device));
    std::vector<int> cnt(R.n);
    HCK(hipMemcpy(cnt.data(), c->nrpy + R.slot0, 4 * (size_t)R.n, hipMemcpyDeviceToHost));
    double e = 0.0;
    for (int v : cnt) e += v;
    out[7] = e;
    if (c->nclp && R.ncl > 0) {
      std::vector<int> cc(R.ncl);
      HCK(hipMemcpy(cc.data(), c->nclp + R.cl0, 4 * (size_t)R.ncl, hipMemcpyDeviceToHost));
      double f = 0.0;
      for (int v : cc) f += v;
      out[8] = f;
    }
  }
  out[0] = R.alpha;
  out[1] = sqrt(R.rc2);
  out[2] = R.skin;
  out[3] = R.K[0];
  out[4] = R.K[1];
  out[5] = R.K[2];
  out[6] = R.order;
  return 0;
}

static int gather_soa(smdlt_ctx* c, const double* h, double* d) {
  // host concatenated (unpadded) -> device slots
  std::vector<double> buf(c->nst, 0.0);
  int k = 0;
  for (int r = 0; r < c->nrep; ++r)
    for (int i = 0; i < c->rh[r].n; ++i) buf[c->rh[r].slot0 + i] = h[k++];
  HCK(hipMemcpyAsync(d, buf.data(), sizeof(double) * c->nst, hipMemcpyHostToDevice, c->stream));
  HCK(hipStreamSynchronize(c->stream));
  return 0;
}
static int scatter_soa(smdlt_ctx* c, const double* d, double* h) {
  std::vector<double> buf(c->nst);
  HCK(hipMemcpyAsync(buf.data(), d, sizeof(double) * c->nst, hipMemcpyDeviceToHost, c->stream));
  HCK(hipStreamSynchronize(c->stream));
  int k = 0;
  for (int r = 0; r < c->nrep; ++r)
    for (int i = 0; i < c->rh[r].n; ++i) h[k++] = buf[c->rh[r].slot0 + i];
  return 0;
}
static int gather_aos(smdlt_ctx* c, const double* h, double* d) {
  std::vector<double> buf(3 * (size_t)c->nst, 0.0);
  size_t k = 0;
  for (int r = 0; r < c->nrep; ++r)
    for (int i = 0; i < c->rh[r].n; ++i)
      for (int q = 0; q < 3; ++q) buf[3 * (size_t)(c->rh[r].slot0 + i) + q] = h[k++];
  HCK(hipMemcpyAsync(d, buf.data(), sizeof(double) * 3 * c->nst, hipMemcpyHostToDevice, c->stream));
  HCK(hipStreamSynchronize(c->stream));
  return 0;
}
static int scatter_aos(smdlt_ctx* c, const double* d, double* h) {
  std::vector<double> buf(3 * (size_t)c->nst);
  HCK(hipMemcpyAsync(buf.data(), d, sizeof(double) * 3 * c->nst, hipMemcpyDeviceToHost, c->stream));
  HCK(hipStreamSynchronize(c->stream));
  size_t k = 0;
  for (int r = 0; r < c->nrep; ++r)
    for (int i = 0; i < c->rh[r].n; ++i)
      for (int q = 0; q < 3; ++q) h[k++] = buf[3 * (size_t)(c->rh[r].slot0 + i) + q];
  return 0;
}

int smdlt_upload_positions(smdlt_ctx* c, const double* rx, const double* ry, const double* rz) {
  if (!c || !rx || !ry || !rz) return SMDLT_E_ARG;
  {  // the remap, cell and brick indices need finite positions
    int n = 0;
    for (int r = 0; r < c->nrep; ++r) n += c->rh[r].n;
    for (int k = 0; k < n; ++k)
      if (!std::isfinite(rx[k]) || !std::isfinite(ry[k]) || !std::isfinite(rz[k])) return SMDLT_E_ARG;
  }
  HCK(hipSetDevice(c->device));
  int st = gather_soa(c, rx, c->rx);
  st = st ? st : gather_soa(c, ry, c->ry);
  st = st ? st : gather_soa(c, rz, c->rz);
  if (st) return st;
  launch_fill_f64(c->stream, c->dxv, 0.0, c->nst);
  launch_fill_f64(c->stream, c->dyv, 0.0, c->nst);
  launch_fill_f64(c->stream, c->dzv, 0.0, c->nst);
  launch_fill_f64(c->stream, c->dsp, 0.0, 3LL * c->nst);
  c->have_pos = true;
  c->lists_built = false;
  c->pos_version++;
  HCK(hipStreamSynchronize(c->stream));
  return 0;
}

int smdlt_download_positions(smdlt_ctx* c, double* rx, double* ry, double* rz) {
  if (!c || !rx || !ry || !rz) return SMDLT_E_ARG;
  HCK(hipSetDevice(c->device));
  int st = scatter_soa(c, c->rx, rx);
  st = st ? st : scatter_soa(c, c->ry, ry);
  st = st ? st : scatter_soa(c, c->rz, rz);
  return st;
}

static int lattice_to_device(smdlt_ctx* c, uint64_t t) {
  c->pos_version++;
  // The lists follow the lattice by one step's strain per neighbour check; a
  // lattice set to a step other than the last one or its successor (a jump in
  // strain) makes the next list use rebuild them.
  if (c->lat_t != UINT64_MAX && t != c->lat_t && t != c->lat_t + 1)
    for (int r = 0; r < c->nrep; ++r)
      if (c->rh[r].edot != 0.0) c->force_rebuild = true;
  c->lat_t = t;
  for (int r = 0; r < c->nrep; ++r) {
    c->ct_prev[r] = c->lat[r].ct;
    c->st_prev[r] = c->lat[r].st;
    c->lat[r].update(t);
    fill_lattice(c, r);
    c->rh[r].reset = (c->lat[r].treal == 0 && t > 0) ? 1 : 0;
    c->rh[r].step = t;
  }
  return upload_reps(c);
}

int smdlt_update_lattice(smdlt_ctx* c, uint64_t t, double* out) {
  if (!c) return SMDLT_E_ARG;
  HCK(hipSetDevice(c->device));
  int st = lattice_to_device(c, t);
  if (st) return st;
  if (out)
    for (int r = 0; r < c->nrep; ++r) {
      const HostLattice& h = c->lat[r];
      double v[8] = {h.L1[0], h.L1[1], h.L2[0], h.L2[1], h.L1xp, h.L2xp, h.L2yp, h.delx};
      memcpy(out + 8 * r, v, sizeof(v));
    }
  HCK(hipStreamSynchronize(c->stream));
  return 0;
}

static int remap_internal(smdlt_ctx* c, uint64_t t) {
  StageScope sc(c, ST_REMAP);
  int st = lattice_to_device(c, t);
  if (st) return st;
  const RepDev& a = c->rh[c->cur_rep_begin];
  const RepDev& b = c->rh[c->cur_rep_end - 1];
  int s0 = a.slot0, ns = b.slot0 + pad_slots(b.n) - s0;
  launch_remap(c->stream, c->d_reps, c->d_slot_rep, s0, ns, c->rx, c->ry, c->rz, c->dxv, c->dyv, c->dzv, c->dsp, 1);
  HCK(hipGetLastError());
  return 0;
}

int smdlt_remap(smdlt_ctx* c, uint64_t t) {
  if (!c) return SMDLT_E_ARG;
  if (!c->have_pos) return SMDLT_E_STATE;
  HCK(hipSetDevice(c->device));
  int st = remap_internal(c, t);
  if (st) return st;
  HCK(hipStreamSynchronize(c->stream));
  return 0;
}

int smdlt_apply_pbc(smdlt_ctx* c) {
  if (!c || !c->have_pos) return SMDLT_E_STATE;
  HCK(hipSetDevice(c->device));
  int st = upload_reps(c);
  if (st) return st;
  launch_remap(c->stream, c->d_reps, c->d_slot_rep, 0, c->nst, c->rx, c->ry, c->rz, c->dxv, c->dyv, c->dzv, c->dsp, 0);
  c->pos_version++;
  HCK(hipStreamSynchronize(c->stream));
  return 0;
}

int smdlt_rotate(smdlt_ctx* c, int inverse) {
  if (!c || !c->have_pos) return SMDLT_E_STATE;
  HCK(hipSetDevice(c->device));
  launch_remap(c->stream, c->d_reps, c->d_slot_rep, 0, c->nst, c->rx, c->ry, c->rz, c->dxv, c->dyv, c->dzv, c->dsp,
               inverse ? 3 : 2);
  c->pos_version++;
  HCK(hipStreamSynchronize(c->stream));
  return 0;
}

int smdlt_min_image(smdlt_ctx* c, int replica, int npairs, const int* i, const int* j, double* out) {
  if (!c || replica < 0 || replica >= c->nrep || npairs < 0 || (npairs && (!i || !j || !out))) return SMDLT_E_ARG;
  if (npairs == 0) return 0;
  HCK(hipSetDevice(c->device));
  for (int k = 0; k < npairs; ++k)
    if (i[k] < 0 || j[k] < 0 || i[k] >= c->rh[replica].n || j[k] >= c->rh[replica].n) return SMDLT_E_ARG;
  int *di, *dj;
  double* dout;
  HCK(hipMalloc(&di, 4 * npairs));
  HCK(hipMalloc(&dj, 4 * npairs));
  HCK(hipMalloc(&dout, 24 * (size_t)npairs));
  HCK(hipMemcpy(di, i, 4 * npairs, hipMemcpyHostToDevice));
  HCK(hipMemcpy(dj, j, 4 * npairs, hipMemcpyHostToDevice));
  launch_min_image(c->stream, c->d_reps + replica, c->rx, c->ry, c->rz, npairs, di, dj, dout);
  HCK(hipStreamSynchronize(c->stream));
  HCK(hipMemcpy(out, dout, 24 * (size_t)npairs, hipMemcpyDeviceToHost));
  hipFree(di);
  hipFree(dj);
  hipFree(dout);
  return 0;
}

// Far-pair table geometry of initLattice and gridHI (S/Initialization.c:247-296,
// X/Initialization.c:258-285, S/gridHI.c:350-398): half-widths from the lattice
// at the end of a period, 2*ceil(half/bin)+1 bins per axis, sizes stretched to fit.
static void table_geometry(const smdlt_ctx* c, int replica, double bin_size, int nb[3], double bs[3]) {
  HostLattice h0;
  const smdlt_params& p = c->P[replica];
  h0.init(p.flow_type, p.L, c->rh[replica].edot, p.dt, p.pbc_compat);
  double xmax, ymax, zmax = p.L / 2;
  if (h0.flow == 2) {
    ymax = -h0.point0[1][1];
    xmax = h0.point0[2][0] * exp(h0.edot * h0.tp * h0.dt);
  } else {
    ymax = p.L / 2;
    int treal = (h0.tp == INT_MAX) ? 0 : h0.tp;
    xmax = p.L / 2 * h0.edot * treal * h0.dt;
  }
  const double half[3] = {xmax, ymax, zmax};
  for (int d = 0; d < 3; ++d) {
    nb[d] = 2 * (int)ceil(half[d] / bin_size) + 1;
    bs[d] = half[d] / ((nb[d] - 1) / 2);
  }
}

int smdlt_bin_index(smdlt_ctx* c, int replica, double bin_size, int npairs, const int* i, const int* j, int* bins) {
  if (!c || replica < 0 || replica >= c->nrep || npairs < 0 || bin_size <= 0) return SMDLT_E_ARG;
  if (npairs == 0) return 0;
  HCK(hipSetDevice(c->device));
  int nbg[3];
  double bs[3];
  table_geometry(c, replica, bin_size, nbg, bs);
  const int nbx = nbg[0], nby = nbg[1], nbz = nbg[2];
  const HostLattice& h = c->lat[replica];
  int nmax = (h.treal <= h.tp / 2) ? 1 : 2;
  int nbv[6] = {nby, nbz, (nbx - 1) / 2, (nby - 1) / 2, (nbz - 1) / 2, nmax};
  int *di, *dj, *dout, *dnb;
  double* dbs;
  HCK(hipMalloc(&di, 4 * npairs));
  HCK(hipMalloc(&dj, 4 * npairs));
  HCK(hipMalloc(&dout, 4 * npairs));
  HCK(hipMalloc(&dnb, sizeof(nbv)));
  HCK(hipMalloc(&dbs, sizeof(bs)));
  HCK(hipMemcpy(di, i, 4 * npairs, hipMemcpyHostToDevice));
  HCK(hipMemcpy(dj, j, 4 * npairs, hipMemcpyHostToDevice));
  HCK(hipMemcpy(dnb, nbv, sizeof(nbv), hipMemcpyHostToDevice));
  HCK(hipMemcpy(dbs, bs, sizeof(bs), hipMemcpyHostToDevice));
  launch_bin_index(c->stream, c->d_reps + replica, c->rx, c->ry, c->rz, npairs, di, dj, dbs, dnb, dout);
  HCK(hipStreamSynchronize(c->stream));
  HCK(hipMemcpy(bins, dout, 4 * npairs, hipMemcpyDeviceToHost));
  hipFree(di);
  hipFree(dj);
  hipFree(dout);
  hipFree(dnb);
  hipFree(dbs);
  return 0;
}

// Device-side neighbour maintenance for the active replica range.
static int neighbors_internal(smdlt_ctx* c, int force) {
  StageScope sc(c, ST_NLIST);
  int rb = c->cur_rep_begin, re = c->cur_rep_end;
  const RepDev& a = c->rh[rb];
  const RepDev& b = c->rh[re - 1];
  int s0 = a.slot0, ns = b.slot0 + pad_slots(b.n) - s0;
  int cb = a.cell0, ce = b.cell0 + c->max_w_cells[re - 1];
  launch_disp_max(c->stream, c->d_reps, rb, re - rb, c->dxv, c->dyv, c->dzv, c->dsp, c->out2);
  if (c->force_rebuild) force = 3;
  c->force_rebuild = false;
  launch_nl_check(c->stream, c->d_reps, rb, re - rb, c->out2, c->aff, force | (c->lists_built ? 0 : 3), c->rep_flags);
  launch_fill_i32(c->stream, c->cell_cnt + cb, 0, ce - cb);
  launch_fill_i32(c->stream, c->cell_fill + cb, 0, ce - cb);
  launch_cell_count(c->stream, c->d_reps, c->d_slot_rep, s0, ns, c->rx, c->ry, c->rz, c->cell_of, c->cell_cnt,
                    c->rep_flags);
  launch_scan(c->stream, c->cell_cnt + cb, c->cell_start + cb, ce - cb);
  launch_cell_fill(c->stream, c->d_reps, c->d_slot_rep, s0, ns, c->cell_of, c->cell_start, c->cell_fill, c->cell_tmp,
                   c->rep_flags);
  launch_cell_sort(c->stream, c->cell_start, c->cell_cnt, cb, ce - cb, c->cell_tmp, c->cell_beads);
  launch_build_lists(c->stream, c->d_reps, c->d_slot_rep, s0, ns, c->rx, c->ry, c->rz, c->cell_start, c->cell_cnt,
                     c->cell_beads, c->rep_flags, c->nst, c->cap_rpy, c->nrpy, c->lrpy, c->cap_lj, c->nlj, c->llj,
                     c->overflow);
  if (c->lcl) {
    const int clb = c->rh[rb].cl0, cle = c->rh[re - 1].cl0 + c->rh[re - 1].ncl;
    launch_build_clusters(c->stream, c->d_reps, c->d_slot_rep, clb, cle - clb, c->cl_words, c->cl_first, c->cl_cnt,
                          c->slot_cl, c->rep_flags, c->nrpy, c->lrpy, c->cap_rpy, c->capc, c->nclp, c->lcl,
                          c->overflow);
  }
  launch_reset_trackers(c->stream, c->d_reps, c->d_slot_rep, s0, ns, c->rep_flags, c->dxv, c->dyv, c->dzv, c->dsp);
  HCK(hipGetLastError());
  c->lists_built = true;
  return 0;
}

static int grow_lists(smdlt_ctx* c) {
  int ov = 0;
  HCK(hipMemcpy(&ov, c->overflow, 4, hipMemcpyDeviceToHost));
  if (!ov) return 0;
  // measure and grow
  c->cap_rpy *= 2;
  c->cap_lj *= 2;
  hipFree(c->lrpy);
  hipFree(c->llj);
  c->lrpy = nullptr;
  c->llj = nullptr;
  if (alloc_dev(&c->lrpy, (size_t)c->cap_rpy * c->nst) || alloc_dev(&c->llj, (size_t)c->cap_lj * c->nst))
    return SMDLT_E_NOMEM;
  if (c->lcl) {
    hipFree(c->lcl);
    c->lcl = nullptr;
    c->capc *= 2;
    if (alloc_dev(&c->lcl, (size_t)c->capc * c->ncl_total + CL_LIST_PAD)) return SMDLT_E_NOMEM;
  }
  HCK(hipMemset(c->overflow, 0, 4));
  destroy_graphs(c);
  return 1;
}

int smdlt_neighbors(smdlt_ctx* c, int force, int* rebuilt) {
  if (!c || !c->have_pos) return SMDLT_E_STATE;
  HCK(hipSetDevice(c->device));
  int st = upload_reps(c);
  if (st) return st;
  for (int attempt = 0; attempt < 8; ++attempt) {
    st = neighbors_internal(c, force ? 3 : 0);
    if (st) return st;
    HCK(hipStreamSynchronize(c->stream));
    int g = grow_lists(c);
    if (g < 0) return g;
    if (g == 0) break;
    force = 1;
  }
  if (rebuilt) {
    std::vector<int> fl(c->nrep);
    HCK(hipMemcpy(fl.data(), c->rep_flags, 4 * c->nrep, hipMemcpyDeviceToHost));
    int n = 0;
    for (int v : fl) n += (v & 1) ? 1 : 0;
    *rebuilt = n;
  }
  return 0;
}

int smdlt_get_verlet(smdlt_ctx* c, int replica, int* nlist, int* list, int stride) {
  if (!c || replica < 0 || replica >= c->nrep || !nlist || !list || stride <= 0) return SMDLT_E_ARG;
  HCK(hipSetDevice(c->device));
  const RepDev& R = c->rh[replica];
  std::vector<int> cnt(R.n), all((size_t)c->cap_lj * c->nst);
  HCK(hipMemcpy(cnt.data(), c->nlj + R.slot0, 4 * R.n, hipMemcpyDeviceToHost));
  HCK(hipMemcpy(all.data(), c->llj, 4 * all.size(), hipMemcpyDeviceToHost));
  for (int i = 0; i < R.n; ++i) {
    int k = 0;
    for (int q = 0; q < cnt[i]; ++q) {
      int j = all[(size_t)q * c->nst + R.slot0 + i] - R.slot0;
      if (j > i) {
        if (k >= stride) return SMDLT_E_ARG;
        list[(size_t)i * stride + k++] = j;
      }
    }
    nlist[i] = k;
  }
  return 0;
}

// The charge displacement log of a step: offsets from the chain centres of mass
// at the step start, then each charge's displacement after the update.
static void charge_rel_internal(smdlt_ctx* c) {
  if (c->nq_total == 0) return;
  launch_charge_rel(c->stream, c->d_reps, c->cur_rep_begin, c->cur_rep_end - c->cur_rep_begin, c->max_chains, c->rx,
                    c->ry, c->rz, c->qrel);
}
static void charge_log_internal(smdlt_ctx* c) {
  if (c->nq_total == 0) return;
  launch_charge_log(c->stream, c->d_reps, c->cur_rep_begin, c->cur_rep_end - c->cur_rep_begin, c->max_nq, c->qtrack,
                    c->mf, c->bz, c->qrel, c->qdisp);
}

// accumulate: a step's forces also add their virial to the stress sums (as
// bondforce/LJforce do every step); a stand-alone force evaluation does not.
static int forces_internal(smdlt_ctx* c, bool accumulate) {
  StageScope sc(c, ST_FORCES);
  int rb = c->cur_rep_begin, re = c->cur_rep_end;
  const RepDev& a = c->rh[rb];
  const RepDev& b = c->rh[re - 1];
  int s0 = a.slot0, ns = b.slot0 + pad_slots(b.n) - s0;
  launch_forces(c->stream, c->d_reps, c->d_slot_rep, s0, ns, c->rx, c->ry, c->rz, c->nst, c->nlj, c->llj, c->f,
                accumulate ? c->vir : nullptr, c->d_err);
  if (accumulate) launch_virial_acc(c->stream, c->d_reps, rb, re - rb, c->vir, c->tau);
  if (c->max_nq > 0)
    launch_coulomb(c->stream, c->d_reps, rb, re - rb, c->rx, c->ry, c->rz, c->qtrack, c->qsign, c->f, c->max_nq);
  HCK(hipGetLastError());
  return 0;
}

static int check_err(smdlt_ctx* c) {
  int e = 0;
  HCK(hipMemcpyAsync(&e, c->d_err, 4, hipMemcpyDeviceToHost, c->stream));
  HCK(hipStreamSynchronize(c->stream));
  if (e) {
    HCK(hipMemset(c->d_err, 0, 4));
    return (e & 1) ? SMDLT_E_FENE_OVERSTRETCH : SMDLT_E_NONFINITE;
  }
  return 0;
}

int smdlt_forces(smdlt_ctx* c, double* fx, double* fy, double* fz) {
  if (!c || !c->have_pos) return SMDLT_E_STATE;
  HCK(hipSetDevice(c->device));
  if (!c->lists_built) {
    int st = smdlt_neighbors(c, 1, nullptr);
    if (st) return st;
  }
  int st = upload_reps(c);
  if (st) return st;
  st = forces_internal(c, false);
  if (st) return st;
  st = check_err(c);
  if (st) return st;
  if (fx || fy || fz) {
    // forces are in the wrap frame; the reference returns lab-frame forces for KRBC
    HCK(hipMemcpyAsync(c->tmp, c->f, 24 * (size_t)c->nst, hipMemcpyDeviceToDevice, c->stream));
    launch_rotate_vec(c->stream, c->d_reps, c->d_slot_rep, 0, c->nst, c->tmp, 1);
    std::vector<double> h(3 * (size_t)smdlt_num_beads(c, -1));
    st = scatter_aos(c, c->tmp, h.data());
    if (st) return st;
    for (size_t i = 0; i < h.size() / 3; ++i) {
      if (fx) fx[i] = h[3 * i];
      if (fy) fy[i] = h[3 * i + 1];
      if (fz) fz[i] = h[3 * i + 2];
    }
  }
  return 0;
}

// u = M v for the active replica range (device pointers).  beta/vout/partial/done
// are the Lanczos hooks (nullptr for a plain product).
static MobArgs mob_args(smdlt_ctx* c, int rb, int re, const double* v, double* u, const double* beta, double* vout,
                        double* partial, const int* done) {
  const RepDev& a = c->rh[rb];
  const RepDev& b = c->rh[re - 1];
  MobArgs m;
  memset(&m, 0, sizeof(m));
  m.reps = c->d_reps;
  m.slot_rep = c->d_slot_rep;
  m.slot_begin = a.slot0;
  m.nslots = b.slot0 + pad_slots(b.n) - a.slot0;
  m.nst = c->nst;
  m.rx = c->rx;
  m.ry = c->ry;
  m.rz = c->rz;
  m.nrpy = c->nrpy;
  m.lrpy = c->lrpy;
  m.cap_rpy = c->cap_rpy;
  m.v = v;
  m.beta = beta;
  m.vout = vout;
  m.u = u;
  m.adot = partial;
  m.done = done;
  m.grid = c->grid;
  m.grid_begin = a.grid0;
  m.grid_len = 0;  // the gather writes every grid point
  m.bsp = c->bsp;
  m.tw = c->tw;
  m.tx = c->tx + c->tx0[rb];
  m.ty = c->ty + c->ty0[rb];
  m.tz = c->tz + c->tz0[rb];
  m.ntx = c->tx0[re] - c->tx0[rb];
  m.nty = c->ty0[re] - c->ty0[rb];
  m.ntz = c->tz0[re] - c->tz0[rb];
  m.ztile = c->ztile;
  m.txy = c->txy + c->txy0[rb];
  m.ntxy = c->txy0[re] - c->txy0[rb];
  m.order = c->rh[rb].order;
  m.bbase = c->bbase;
  m.bw = c->bw;
  m.brick_cnt = c->brick_cnt;
  m.brick_start = c->brick_start;
  m.brick_fill = c->brick_fill;
  m.brick_beads = c->brick_beads;
  m.bricks = c->bricks + c->bk0[rb];
  m.nbricks = c->bk0[re] - c->bk0[rb];
  m.brick_begin = c->bk0[rb];
  m.brick_len = c->bk0[re] - c->bk0[rb];
  m.p32 = c->p32;
  m.v32 = c->v32;
  m.bcand_cnt = c->bcand_cnt;
  m.bcand_start = c->bcand_start;
  m.bcand = c->bcand;
  m.rtab = c->rtab;
  m.cl_first = c->cl_first;
  m.cl_cnt = c->cl_cnt;
  m.lcl = c->lcl;
  m.nclp = c->nclp;
  m.capc = c->capc;
  m.cl_begin = a.cl0;
  m.ncl = b.cl0 + b.ncl - a.cl0;
  m.ntab_max = 0;
  for (int r = rb; r < re; ++r) m.ntab_max = std::max(m.ntab_max, c->rh[r].ntab);
  c->fch_buf.clear();
  auto bound = [&](int r) {
    c->fch_buf.push_back(c->tx0[r] - c->tx0[rb]);
    c->fch_buf.push_back(c->ty0[r] - c->ty0[rb]);
    c->fch_buf.push_back(c->tz0[r] - c->tz0[rb]);
    c->fch_buf.push_back(c->txy0[r] - c->txy0[rb]);
  };
  bound(rb);
  for (int r : c->fch_rep)
    if (r > rb && r < re) bound(r);
  bound(re);
  m.fch = c->fch_buf.data();
  m.nfch = (int)c->fch_buf.size() / 4 - 1;
  return m;
}

static void mob_launch(smdlt_ctx* c, MobArgs& m) {
  hipEvent_t ev[9];
  std::vector<hipEvent_t> evc;
  const int nch = m.nfch > 1 ? m.nfch : 1;
  m.ev = nullptr;
  m.evc = nullptr;
  if (c->prof) {
    for (int k = 0; k < 9; ++k) ev[k] = get_event(c);
    m.ev = ev;
    if (nch > 1) {
      evc.resize(6 * (size_t)nch);
      for (auto& e : evc) e = get_event(c);
      m.evc = evc.data();
    }
  }
  if (c->prec == SMDLT_PREC_FP64)
    launch_mobility<double>(c->stream, m);
  else
    launch_mobility<float>(c->stream, m);
  if (m.ev) {  // kernel k runs between ev[k] and ev[k+1]; each event is returned once
    for (int k = 0; k < 8; ++k) {
      const bool fft = k >= 2 && k <= 6;  // with FFT chunks, the chunks' own events time the passes
      c->timers.push_back({ev[k], ev[k + 1], (nch > 1 && fft) ? -1 : ST_K_ZERO + k, k == 0});
    }
    for (int q = 0; q < (int)evc.size() / 6; ++q)
      for (int k = 0; k < 5; ++k)
        c->timers.push_back({evc[6 * q + k], evc[6 * q + k + 1], ST_K_ZERO + 2 + k, k == 0, q == 0 ? 1 : 0});
  }
}

// One launch sequence for every replica of the range, the grid passes in FFT
// chunks.  (Running the whole sequence in replica groups whose grids fit the
// 256 MB last-level cache was measured slower on 64 replicas, 380k beads: the
// smaller launches lose more to tails than the grid passes gain from the cache.)
static void mobility_internal(smdlt_ctx* c, const double* v, double* u, const double* beta, double* vout,
                              double* partial, const int* done) {
  StageScope sc(c, ST_MOB);
  const int rb = c->cur_rep_begin, re = c->cur_rep_end;
  MobArgs m = mob_args(c, rb, re, v, u, beta, vout, partial, done);
  m.prep = (c->prep_version != c->pos_version || c->prep_range != rb * 100003LL + re) ? 1 : 0;
  c->prep_version = c->pos_version;
  c->prep_range = rb * 100003LL + re;
  m.phase = 3;
  mob_launch(c, m);
}

static int ensure_lists(smdlt_ctx* c) {
  if (c->lists_built && !c->force_rebuild) return 0;
  int st = smdlt_neighbors(c, 1, nullptr);
  return st;
}

int smdlt_mobility_apply_device(smdlt_ctx* c, const double* d_v, double* d_u) {
  if (!c || !d_v || !d_u) return SMDLT_E_ARG;
  if (!c->have_pos) return SMDLT_E_STATE;
  HCK(hipSetDevice(c->device));
  int st = ensure_lists(c);
  if (st) return st;
  mobility_internal(c, d_v, d_u, nullptr, nullptr, nullptr, nullptr);
  HCK(hipGetLastError());
  return 0;
}

int smdlt_mobility_apply(smdlt_ctx* c, const double* v, double* u) {
  if (!c || !v || !u) return SMDLT_E_ARG;
  if (!c->have_pos) return SMDLT_E_STATE;
  HCK(hipSetDevice(c->device));
  int st = ensure_lists(c);
  if (st) return st;
  st = upload_reps(c);
  if (st) return st;
  st = gather_aos(c, v, c->tmp);
  if (st) return st;
  mobility_internal(c, c->tmp, c->tmp2, nullptr, nullptr, nullptr, nullptr);
  HCK(hipGetLastError());
  st = scatter_aos(c, c->tmp2, u);
  collect_timers(c);
  return st;
}

// Lanczos on the active replica range: y = M^(1/2) z (device pointers).
// Lanczos pieces: start (|z|, flags), iteration j, combination y = V c.
static void lz_start(smdlt_ctx* c, const double* z) {
  int rb = c->cur_rep_begin, re = c->cur_rep_end;
  int s0 = c->rh[rb].slot0, ns = c->rh[re - 1].slot0 + pad_slots(c->rh[re - 1].n) - s0;
  StageScope sc(c, ST_LANCZOS);
  launch_fill_i32(c->stream, c->done + rb, 0, re - rb);
  lz_reset_checks(c->stream, c->scal, rb, re - rb);
  lz_wave_dot(c->stream, c->d_reps, c->d_slot_rep, s0, ns, z, nullptr, nullptr, c->partial);
  lz_rep_reduce(c->stream, c->d_reps, rb, re - rb, c->partial, nullptr, c->scal, SCAL_STRIDE, 4 * LZ_MAXIT, 1,
                c->bdiv);
}

static void lz_iteration(smdlt_ctx* c, const double* z, int j, double tol, int maxit) {
  int rb = c->cur_rep_begin, re = c->cur_rep_end;
  int s0 = c->rh[rb].slot0, ns = c->rh[re - 1].slot0 + pad_slots(c->rh[re - 1].n) - s0;
  size_t n3 = 3 * (size_t)c->nst;
  const double* vin = (j == 1) ? z : c->W;
  double* vj = c->V + (size_t)(j - 1) * n3;
  const double* vjm1 = (j > 1) ? c->V + (size_t)(j - 2) * n3 : nullptr;
  // the cluster pass (mixed precision) also forms alpha's partials, one per cluster
  const bool fused_dot = c->lcl != nullptr && c->prec == SMDLT_PREC_MIXED;
  mobility_internal(c, vin, c->W, c->bdiv, vj, fused_dot ? c->adot_cl : nullptr, c->done);
  StageScope sc(c, ST_LANCZOS);
  if (!fused_dot) lz_wave_dot(c->stream, c->d_reps, c->d_slot_rep, s0, ns, vj, c->W, c->done, c->partial2);
  lz_orth(c->stream, c->d_reps, c->d_slot_rep, s0, ns, c->W, vj, vjm1, c->scal, j, c->done,
          fused_dot ? nullptr : c->partial2, fused_dot ? c->adot_cl : nullptr, c->partial);
  lz_tri(c->stream, c->d_reps, rb, re - rb, c->partial, c->scal, j, tol, maxit, c->done, c->bdiv, c->iters);
}

static void lz_finish(smdlt_ctx* c, double* y) {
  int rb = c->cur_rep_begin, re = c->cur_rep_end;
  int s0 = c->rh[rb].slot0, ns = c->rh[re - 1].slot0 + pad_slots(c->rh[re - 1].n) - s0;
  StageScope sc(c, ST_LANCZOS);
  lz_combine(c->stream, c->d_reps, c->d_slot_rep, s0, ns, c->V, c->nst, c->scal, c->iters, y);
}

// copy done flags and iteration counts of the current replica range to the
// host; returns 1 when every replica has stopped
static int lz_poll(smdlt_ctx* c, int* max_iters, int* overflow = nullptr) {
  int rb = c->cur_rep_begin, re = c->cur_rep_end, nr = re - rb;
  HCK(hipMemcpyAsync(c->h_flags, c->done + rb, 4 * nr, hipMemcpyDeviceToHost, c->stream));
  HCK(hipMemcpyAsync(c->h_flags + c->nrep, c->iters + rb, 4 * nr, hipMemcpyDeviceToHost, c->stream));
  if (overflow) HCK(hipMemcpyAsync(c->h_flags + 2 * c->nrep, c->overflow, 4, hipMemcpyDeviceToHost, c->stream));
  HCK(hipStreamSynchronize(c->stream));
  if (overflow) *overflow = c->h_flags[2 * c->nrep];
  bool all = true;
  int mx = 0;
  for (int r = 0; r < nr; ++r) {
    all = all && c->h_flags[r] != 0;
    mx = std::max(mx, c->h_flags[c->nrep + r]);
  }
  if (max_iters) *max_iters = mx;
  return all ? 1 : 0;
}

static int brownian_internal(smdlt_ctx* c, const double* z, double* y, double tol, int maxit, bool host_exit) {
  lz_start(c, z);
  for (int j = 1; j <= maxit; ++j) {
    lz_iteration(c, z, j, tol, maxit);
    if (host_exit && j >= 2) {
      int all = lz_poll(c, nullptr);
      if (all < 0) return all;
      if (all) break;
    }
  }
  lz_finish(c, y);
  HCK(hipGetLastError());
  return 0;
}

static int read_iters(smdlt_ctx* c, int* iters, double* resid) {
  HCK(hipMemcpyAsync(c->last_iters.data(), c->iters, 4 * c->nrep, hipMemcpyDeviceToHost, c->stream));
  std::vector<double> sc((size_t)SCAL_STRIDE * c->nrep);
  HCK(hipMemcpyAsync(sc.data(), c->scal, 8 * sc.size(), hipMemcpyDeviceToHost, c->stream));
  std::vector<int> dn(c->nrep);
  HCK(hipMemcpyAsync(dn.data(), c->done, 4 * c->nrep, hipMemcpyDeviceToHost, c->stream));
  HCK(hipStreamSynchronize(c->stream));
  int st = 0;
  for (int r = c->cur_rep_begin; r < c->cur_rep_end; ++r) {
    if (iters) iters[r] = c->last_iters[r];
    if (resid) resid[r] = sc[(size_t)r * SCAL_STRIDE + 4 * LZ_MAXIT + 2];
    if (dn[r] != 1) st = SMDLT_E_LANCZOS_NOCONV;
  }
  return st;
}

int smdlt_brownian(smdlt_ctx* c, const double* z, double* y, double tol, int maxit, int* iters, double* resid) {
  if (!c || !z || !y) return SMDLT_E_ARG;
  if (!c->have_pos) return SMDLT_E_STATE;
  if (maxit <= 0) maxit = c->maxit;
  if (maxit > c->maxit) return SMDLT_E_ARG;
  if (tol <= 0) tol = 1e-6;
  HCK(hipSetDevice(c->device));
  int st = ensure_lists(c);
  if (st) return st;
  st = upload_reps(c);
  if (st) return st;
  st = gather_aos(c, z, c->z);
  if (st) return st;
  st = brownian_internal(c, c->z, c->bz, tol, maxit, true);
  if (st) return st;
  int conv = read_iters(c, iters, resid);
  st = scatter_aos(c, c->bz, y);
  collect_timers(c);
  return st ? st : conv;
}

int smdlt_lanczos_status(smdlt_ctx* c, int* iters, int* status, double* change) {
  if (!c) return SMDLT_E_ARG;
  HCK(hipSetDevice(c->device));
  std::vector<int> it(c->nrep), dn(c->nrep);
  std::vector<double> sc((size_t)SCAL_STRIDE * c->nrep);
  HCK(hipMemcpyAsync(it.data(), c->iters, 4 * c->nrep, hipMemcpyDeviceToHost, c->stream));
  HCK(hipMemcpyAsync(dn.data(), c->done, 4 * c->nrep, hipMemcpyDeviceToHost, c->stream));
  HCK(hipMemcpyAsync(sc.data(), c->scal, 8 * sc.size(), hipMemcpyDeviceToHost, c->stream));
  HCK(hipStreamSynchronize(c->stream));
  for (int r = 0; r < c->nrep; ++r) {
    if (iters) iters[r] = it[r];
    if (status) status[r] = dn[r];
    if (change) change[r] = sc[(size_t)r * SCAL_STRIDE + 4 * LZ_MAXIT + 2];
  }
  return 0;
}

int smdlt_last_iters(const smdlt_ctx* c, int* iters) {
  if (!c || !iters) return SMDLT_E_ARG;
  for (int r = 0; r < c->nrep; ++r) iters[r] = c->last_iters[r];
  return 0;
}

// Periodic self-mobility of one bead for the replica's current lattice and
// split: selfconst + sum_k G(k)_aa (the Ewald sum of its own images).
static void self_mobility(const RepDev& R, double out[3]) {
  double a2 = R.alpha * R.alpha;
  double kmax = choose_kmax(R.alpha, 1e-14);
  out[0] = out[1] = out[2] = R.selfconst;
  // |m1| <= kmax/|b1x|; |m2 b2y| <= kmax + |m1 b1y|
  int m1m = (int)ceil(kmax / fabs(R.kb[0])) + 1;
  int m2m = (int)ceil((kmax + m1m * fabs(R.kb[1])) / fabs(R.kb[2])) + 1;
  int m3m = (int)ceil(kmax / R.kb[3]) + 1;
  for (int m1 = -m1m; m1 <= m1m; ++m1)
    for (int m2 = -m2m; m2 <= m2m; ++m2)
      for (int m3 = -m3m; m3 <= m3m; ++m3) {
        if (!m1 && !m2 && !m3) continue;
        double kx = m1 * R.kb[0], ky = m1 * R.kb[1] + m2 * R.kb[2], kz = m3 * R.kb[3];
        double kk = kx * kx + ky * ky + kz * kz;
        if (kk > kmax * kmax) continue;
        double q = kk / (4 * a2);
        double g = (1 - kk / 3) * (1 + q + 2 * q * q) * 6 * M_PI / kk * exp(-q) / R.vol;
        out[0] += g * (1 - kx * kx / kk);
        out[1] += g * (1 - ky * ky / kk);
        out[2] += g * (1 - kz * kz / kk);
      }
}

int smdlt_tea_apply(smdlt_ctx* c, const double* beta, const double* C, const double* z, double* y) {
  if (!c || !beta || !C || !z || !y) return SMDLT_E_ARG;
  if (!c->have_pos) return SMDLT_E_STATE;
  HCK(hipSetDevice(c->device));
  int st = ensure_lists(c);
  if (st) return st;
  st = upload_reps(c);
  if (st) return st;
  std::vector<double> diag(3 * c->nrep);
  for (int r = 0; r < c->nrep; ++r) self_mobility(c->rh[r], &diag[3 * r]);
  double *dd, *db;
  HCK(hipMalloc(&dd, 8 * diag.size()));
  HCK(hipMalloc(&db, 8 * c->nrep));
  HCK(hipMemcpy(dd, diag.data(), 8 * diag.size(), hipMemcpyHostToDevice));
  HCK(hipMemcpy(db, beta, 8 * c->nrep, hipMemcpyHostToDevice));
  st = gather_aos(c, z, c->z);
  if (!st) st = gather_aos(c, C, c->bz);
  if (st) return st;
  mobility_internal(c, c->z, c->tmp2, nullptr, nullptr, nullptr, nullptr);
  lz_tea(c->stream, c->d_slot_rep, 0, c->nst, c->tmp2, c->z, dd, db, c->bz, c->tmp);
  HCK(hipGetLastError());
  st = scatter_aos(c, c->tmp, y);
  hipFree(dd);
  hipFree(db);
  return st;
}

int smdlt_set_tea(smdlt_ctx* c, const double* beta, const double* C) {
  if (!c || !beta || !C) return SMDLT_E_ARG;
  HCK(hipSetDevice(c->device));
  if (!c->tea_beta && (alloc_dev(&c->tea_beta, c->nrep) || alloc_dev(&c->tea_C, 3 * (size_t)c->nst) ||
                       alloc_dev(&c->tea_diag, 3 * (size_t)c->nrep)))
    return SMDLT_E_NOMEM;
  std::vector<double> diag(3 * c->nrep);
  for (int r = 0; r < c->nrep; ++r) self_mobility(c->rh[r], &diag[3 * r]);
  HCK(hipMemcpy(c->tea_diag, diag.data(), 8 * diag.size(), hipMemcpyHostToDevice));
  HCK(hipMemcpy(c->tea_beta, beta, 8 * (size_t)c->nrep, hipMemcpyHostToDevice));
  return gather_aos(c, C, c->tea_C);
}

// ---------------------------------------------------------------- TEA sample
// The reciprocal sum of ewaldBin (S/main26.c:1000-1097, X/main14.c:1034-1131) as
// the reference bounds it: for kx = 0..kxmax the ky range runs from the first to
// the last in-cut value (first exit ends it), likewise kz within each (kx, ky);
// kxmax is then fixed by the reference's while-loop; M2(k) per vector, the kx = 0
// plane halved and k = 0 removed.  Shear names: rkx from ky, rky from kx; KRBC:
// kcoeff = 2 pi/detL over the deformed lattice.  Appends 12 doubles per vector
// (rkx, rky, rkz, M2[9]) and returns the count.
static int tea_kspace(const HostLattice& h, double M_HI, double rce, double vol, std::vector<double>& tab) {
  const double L = h.L, detL = h.detL, L2yp = h.L2yp;
  const double* L1 = h.L1;
  const double* L2 = h.L2;
  const bool krbc = h.flow == 2;
  const double alpha = M_HI / rce, a2 = alpha * alpha, a4 = a2 * a2;
  const double rkcut = 2.0 * M_HI * M_HI / rce, rkkcut = rkcut * rkcut;
  int kxmax = (int)ceil(rkcut * detL / (M_PI * 2.0 * L * L2[1]));
  const int kymax = (int)ceil(rkcut * detL / (M_PI * 2.0 * L * L1[0]));
  const int kzmax = (int)ceil(rkcut * L / (M_PI * 2.0));
  const int kxalloc = kxmax + 1, ny = 2 * kymax + 1;
  std::vector<int> kyl(kxalloc + 2, 0), kyu(kxalloc + 2, 0), kzl((size_t)(kxalloc + 2) * ny, 0),
      kzu((size_t)(kxalloc + 2) * ny, 0);
  const double kc = krbc ? 2.0 * M_PI / detL : 2.0 * M_PI;
  auto rk = [&](int kx, int ky, double& rkx, double& rky) {
    if (krbc) {
      rkx = kc * L * (kx * L2[1] - ky * L1[1]);
      rky = kc * L * (-kx * L2[0] + ky * L1[0]);
    } else {
      rky = kc * (kx - ky * (L2[0] / L1[0])) / L2[1];
      rkx = kc * ky / L1[0];
    }
  };
  for (int kx = 0; kx <= kxmax; ++kx) {
    bool ylo = false, yhi = false;
    for (int ky = -kymax; ky <= kymax; ++ky) {
      double rkx, rky;
      rk(kx, ky, rkx, rky);
      const double rkxy = rkx * rkx + rky * rky;
      bool zlo = false, zhi = false;
      for (int kz = -kzmax; kz <= kzmax; ++kz) {
        const double rkz = kz * M_PI * 2.0 / L, rkk = rkxy + rkz * rkz;
        if (rkk < rkkcut && !zlo) {
          kzl[(size_t)ny * kx + ky + kymax] = kz;
          zlo = true;
        }
        if (rkk > rkkcut && zlo && !zhi) {
          kzu[(size_t)ny * kx + ky + kymax] = kz - 1;
          zhi = true;
        }
      }
      if (rkxy < rkkcut && !ylo) {
        kyl[kx] = ky;
        ylo = true;
      }
      if (rkxy > rkkcut && ylo && !yhi) {
        kyu[kx] = ky - 1;
        yhi = true;
      }
    }
  }
  int kx = 0;
  while (kx < kxalloc && kyu[kx] != -kymax && kx <= (int)(rkcut * detL / (M_PI * 2.0 * L * L2yp))) {
    kxmax = kx;
    ++kx;
  }
  kxmax = std::min(kxmax + 1, kxalloc - 1);
  int count = 0;
  for (int kx = 0; kx <= kxmax; ++kx)
    for (int ky = kyl[kx]; ky <= kyu[kx]; ++ky) {
      double rkx, rky;
      rk(kx, ky, rkx, rky);
      const double rkxy = rkx * rkx + rky * rky;
      for (int kz = kzl[(size_t)ny * kx + ky + kymax]; kz <= kzu[(size_t)ny * kx + ky + kymax]; ++kz) {
        const double rkz = kz * M_PI * 2.0 / L, rkk = rkxy + rkz * rkz;
        const double m2 = 2.0 * (1.0 - rkk / 3.0) * (1.0 + rkk / (4.0 * a2) + rkk * rkk / (8.0 * a4)) * 6.0 * M_PI /
                          rkk * exp(-rkk / (4.0 * a2));
        double m[9] = {m2 * (1 - rkx * rkx / rkk) / vol,  m2 * (-rkx * rky / rkk) / vol, m2 * (-rkx * rkz / rkk) / vol,
                       m2 * (-rky * rkx / rkk) / vol,     m2 * (1 - rky * rky / rkk) / vol, m2 * (-rky * rkz / rkk) / vol,
                       m2 * (-rkz * rkx / rkk) / vol,     m2 * (-rkz * rky / rkk) / vol, m2 * (1 - rkz * rkz / rkk) / vol};
        if (kx == 0)
          for (double& v : m) v = (ky == 0 && kz == 0) ? 0.0 : v / 2.0;
        tab.push_back(rkx);
        tab.push_back(rky);
        tab.push_back(rkz);
        for (double v : m) tab.push_back(v);
        ++count;
      }
    }
  return count;
}

// The reference's Ewald split for one lattice state (ewaldBin M_HI = 3.5, gridHI
// M_HI = 6): rce and the image range over the period, alpha powers, lattice
// vectors, its k table appended to tab, and the self block Dtr with its quirks
// (S/main26.c:1097-1130: images with all three indices nonzero, z entry with
// dy*dy, no rce test).
static TeaRep tea_rep_for(const HostLattice& h, double M_HI, std::vector<double>& tab) {
  TeaRep t{};
  const bool first_half = h.treal <= h.tp / 2;
  t.rce = first_half ? h.L2yp : 2.0 * h.L2yp;
  t.nmax = first_half ? 1 : 2;
  t.alpha = M_HI / t.rce;
  t.alpha2 = t.alpha * t.alpha;
  t.alpha3 = t.alpha2 * t.alpha;
  t.alpha5 = t.alpha3 * t.alpha2;
  t.alpha7 = t.alpha5 * t.alpha2;
  t.L1[0] = h.L1[0];
  t.L1[1] = h.L1[1];
  t.L2[0] = h.L2[0];
  t.L2[1] = h.L2[1];
  t.L = h.L;
  t.k0 = (int)(tab.size() / 12);
  t.nk = tea_kspace(h, M_HI, t.rce, h.L * h.L * h.L, tab);
  const double a = t.alpha, rtpi = sqrt(M_PI);
  double sc = 1 - 6 / rtpi * a + 40 / (3 * rtpi) * t.alpha3;
  t.dtr[0] = t.dtr[1] = t.dtr[2] = sc;
  for (int q = 0; q < t.nk; ++q) {
    const double* m = &tab[12 * (size_t)(t.k0 + q) + 3];
    t.dtr[0] += m[0];
    t.dtr[1] += m[4];
    t.dtr[2] += m[8];
  }
  for (int nx = -t.nmax; nx <= t.nmax; ++nx)
    for (int ny = -t.nmax; ny <= t.nmax; ++ny)
      for (int nz = -t.nmax; nz <= t.nmax; ++nz) {
        if (!(nx != 0 && ny != 0 && nz != 0)) continue;
        double dx = nx * h.L1[0] + ny * h.L2[0], dy = nx * h.L1[1] + ny * h.L2[1], dz = nz * h.L;
        double rr = dx * dx + dy * dy + dz * dz, rd = sqrt(rr);
        double C1 = 0.75 / rd + 0.5 / (rd * rr);
        double C2 =
            4.0 * t.alpha7 * rr * rr + 3.0 * t.alpha3 * rr - 20.0 * t.alpha5 * rr - 4.5 * a + 14.0 * t.alpha3 + a / rr;
        double C3 = 0.75 / rd - 1.5 / (rd * rr);
        double C4 = -4.0 * t.alpha7 * rr * rr - 3.0 * t.alpha3 * rr + 16.0 * t.alpha5 * rr + 1.5 * a -
                    2.0 * t.alpha3 - 3.0 * a / rr;
        double term1 = C1 * erfc(a * rd) + C2 * exp(-t.alpha2 * rr) / rtpi;
        double term2 = C3 * erfc(a * rd) + C4 * exp(-t.alpha2 * rr) / rtpi;
        t.dtr[0] += term1 + term2 * dx * dx / rr;
        t.dtr[1] += term1 + term2 * dy * dy / rr;
        t.dtr[2] += term1 + term2 * dy * dy / rr;
      }
  return t;
}

int smdlt_tea_sample(smdlt_ctx* c, double* beta, double* C, double* eps) {
  if (!c || !beta || !C) return SMDLT_E_ARG;
  if (!c->have_pos) return SMDLT_E_STATE;
  HCK(hipSetDevice(c->device));
  const int rb = c->cur_rep_begin, re = c->cur_rep_end;
  std::vector<TeaRep> tr;
  std::vector<double> tab;
  std::vector<int> brep;
  const double M_HI = 3.5;
  for (int r = rb; r < re; ++r) {
    const HostLattice& h = c->lat[r];
    const RepDev& R = c->rh[r];
    TeaRep t = tea_rep_for(h, M_HI, tab);
    if (t.nk > TEA_MAXK) return SMDLT_E_GEOMETRY;
    t.slot0 = R.slot0;
    t.n = R.n;
    t.bead0 = (int)brep.size();
    for (int i = 0; i < R.n; ++i) brep.push_back(r - rb);
    tr.push_back(t);
  }
  const int nb = (int)brep.size();
  TeaRep* d_tr = nullptr;
  int* d_brep = nullptr;
  double *d_tab = nullptr, *d_eps = nullptr, *d_rs = nullptr;
  int st = alloc_dev(&d_tr, tr.size());
  st |= alloc_dev(&d_brep, brep.size());
  st |= alloc_dev(&d_tab, tab.size());
  st |= alloc_dev(&d_eps, (size_t)c->nst);
  st |= alloc_dev(&d_rs, 3 * (size_t)c->nst);
  std::vector<double> he(c->nst), hr(3 * (size_t)c->nst);
  if (!st) {
    HCK(hipMemcpy(d_tr, tr.data(), sizeof(TeaRep) * tr.size(), hipMemcpyHostToDevice));
    HCK(hipMemcpy(d_brep, brep.data(), sizeof(int) * brep.size(), hipMemcpyHostToDevice));
    if (!tab.empty()) HCK(hipMemcpy(d_tab, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice));
    launch_tea_rows(c->stream, d_tr, d_brep, nb, c->rx, c->ry, c->rz, d_tab, d_eps, d_rs);
    HCK(hipGetLastError());
    HCK(hipMemcpyAsync(he.data(), d_eps, 8 * he.size(), hipMemcpyDeviceToHost, c->stream));
    HCK(hipMemcpyAsync(hr.data(), d_rs, 8 * hr.size(), hipMemcpyDeviceToHost, c->stream));
    HCK(hipStreamSynchronize(c->stream));
  }
  for (void* p : {(void*)d_tr, (void*)d_brep, (void*)d_tab, (void*)d_eps, (void*)d_rs})
    if (p) hipFree(p);
  if (st) return SMDLT_E_NOMEM;
  // CATEA (S/main26.c:1317-1381) per replica: eps/(9N^2 - 3N) (in FP64: the
  // reference's int expression overflows above N = 15,447), beta, C
  size_t off = 0;
  for (int r = 0; r < c->nrep; ++r) {
    const RepDev& R = c->rh[r];
    if (r >= rb && r < re) {
      const double N = R.n, ROW = 3.0 * R.n;
      double es = 0.0;
      for (int i = 0; i < R.n; ++i) es += he[R.slot0 + i];
      const double e = es / (9.0 * N * N - 3.0 * N), e2 = e * e;
      const double b = (1.0 - sqrt(1.0 - (ROW * e2 - ROW * e))) / (ROW * e2 - ROW * e);
      beta[r] = b;
      if (eps) eps[r] = e;
      for (int i = 0; i < 3 * R.n; ++i) C[off + i] = sqrt(1.0 / (1.0 + hr[3 * (size_t)R.slot0 + i] * b * b));
    }
    off += 3 * (size_t)R.n;
  }
  return 0;
}

int smdlt_far_table_dims(smdlt_ctx* c, int replica, double bin_size, int dims[3], double sizes[3]) {
  if (!c || !dims || replica < 0 || replica >= c->nrep || bin_size <= 0) return SMDLT_E_ARG;
  if (c->P[replica].flow_type == SMDLT_FLOW_EQUILIBRIUM || c->rh[replica].edot == 0.0) return SMDLT_E_ARG;
  double bs[3];
  table_geometry(c, replica, bin_size, dims, bs);
  if (sizes)
    for (int d = 0; d < 3; ++d) sizes[d] = bs[d];
  return 0;
}

int smdlt_far_table(smdlt_ctx* c, int replica, int nt_avg, double bin_size, double* table) {
  int nb[3];
  double bs[3];
  int st = smdlt_far_table_dims(c, replica, bin_size, nb, bs);
  if (st) return st;
  if (!table || nt_avg <= 0) return SMDLT_E_ARG;
  HCK(hipSetDevice(c->device));
  const smdlt_params& p = c->P[replica];
  HostLattice h;
  h.init(p.flow_type, p.L, c->rh[replica].edot, p.dt, p.pbc_compat);
  std::vector<TeaRep> ph;
  std::vector<double> tab;
  for (int nt = 0; nt < nt_avg; ++nt) {  // t = nt*tp/nt_avg in the reference's int arithmetic
    h.update((uint64_t)((long long)nt * h.tp / nt_avg));
    ph.push_back(tea_rep_for(h, 6.0, tab));
  }
  const size_t nbins = (size_t)nb[0] * nb[1] * nb[2], n = 9 * nbins * nt_avg;
  TeaRep* d_ph = nullptr;
  double *d_tab = nullptr, *d_out = nullptr;
  st = alloc_dev(&d_ph, ph.size());
  st |= alloc_dev(&d_tab, tab.size());
  st |= alloc_dev(&d_out, n);
  if (!st) {
    HCK(hipMemcpy(d_ph, ph.data(), sizeof(TeaRep) * ph.size(), hipMemcpyHostToDevice));
    if (!tab.empty()) HCK(hipMemcpy(d_tab, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice));
    launch_far_table(c->stream, d_ph, nt_avg, d_tab, nb, bs, d_out);
    HCK(hipGetLastError());
    HCK(hipMemcpyAsync(table, d_out, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    HCK(hipStreamSynchronize(c->stream));
  }
  for (void* q : {(void*)d_ph, (void*)d_tab, (void*)d_out})
    if (q) hipFree(q);
  return st ? SMDLT_E_NOMEM : 0;
}

int smdlt_write_far_table(const char* path, const double* table, int nt_avg, const int dims[3]) {
  if (!path || !table || !dims || nt_avg <= 0) return SMDLT_E_ARG;
  FILE* f = fopen(path, "w");
  if (!f) return SMDLT_E_ARG;
  const size_t nbins = (size_t)dims[0] * dims[1] * dims[2];
  for (int nt = 0; nt < nt_avg; ++nt)
    for (int i = 0; i < dims[0]; ++i)
      for (int j = 0; j < dims[1]; ++j)
        for (int k = 0; k < dims[2]; ++k) {
          const double* D = table + 9 * (nt * nbins + ((size_t)i * dims[1] + j) * dims[2] + k);
          fprintf(f, "%d %d %d 1\n", i, j, k);
          for (int a = 0; a < 3; ++a) {
            for (int b = 0; b < 3; ++b) fprintf(f, "%.6e ", D[3 * a + b]);
            fprintf(f, "\n");
          }
        }
  return fclose(f) == 0 ? 0 : SMDLT_E_ARG;
}

// One timestep for the active replica range; noise_dev: device noise (nullptr =
// generate).  Lanczos early exit on the host only when host_exit.
static int step_internal(smdlt_ctx* c, uint64_t t, double* noise_dev, bool host_exit) {
  int rb = c->cur_rep_begin, re = c->cur_rep_end;
  const RepDev& a = c->rh[rb];
  const RepDev& b = c->rh[re - 1];
  int s0 = a.slot0, ns = b.slot0 + pad_slots(b.n) - s0;
  const smdlt_params& p = c->P[rb];
  if (p.noise_mode == SMDLT_NOISE_TEA && !c->tea_beta) return SMDLT_E_STATE;  // smdlt_set_tea first
  int st = remap_internal(c, t);
  if (st) return st;
  // host noise is the reference's lab-frame R; rotate it into the wrap frame
  if (noise_dev) launch_rotate_vec(c->stream, c->d_reps, c->d_slot_rep, s0, ns, noise_dev, 0);
  st = neighbors_internal(c, 0);
  if (st) return st;
  if (host_exit) {  // grow the lists now rather than finish the step with truncated ones
    HCK(hipStreamSynchronize(c->stream));
    for (int attempt = 0; attempt < 8; ++attempt) {
      int g = grow_lists(c);
      if (g < 0) return g;
      if (g == 0) break;
      st = neighbors_internal(c, 3);
      if (st) return st;
      HCK(hipStreamSynchronize(c->stream));
    }
  }
  st = forces_internal(c, true);
  if (st) return st;
  charge_rel_internal(c);
  const double* zz = noise_dev;
  {
    StageScope sc(c, ST_NOISE);
    if (!zz) {
      launch_noise(c->stream, c->d_slot_rep, c->d_reps, s0, ns, c->d_seeds, t, c->z);
      zz = c->z;
    }
  }
  if (p.noise_mode == SMDLT_NOISE_FREE_DRAINING) {  // updateFD: M = I (S/main26.c:731-744)
    launch_copy_f64(c->stream, c->mf, c->f, 3LL * c->nst);
    launch_copy_f64(c->stream, c->bz, zz, 3LL * c->nst);
  } else if (p.noise_mode == SMDLT_NOISE_TEA) {  // updateHI's TEA product (S/main26.c:787-804)
    mobility_internal(c, c->f, c->mf, nullptr, nullptr, nullptr, nullptr);
    mobility_internal(c, zz, c->tmp2, nullptr, nullptr, nullptr, nullptr);
    lz_tea(c->stream, c->d_slot_rep, s0, ns, c->tmp2, zz, c->tea_diag, c->tea_beta, c->tea_C, c->bz);
  } else {
    mobility_internal(c, c->f, c->mf, nullptr, nullptr, nullptr, nullptr);
    st = brownian_internal(c, zz, c->bz, p.lanczos_tol > 0 ? p.lanczos_tol : 1e-6, c->maxit, host_exit);
    if (st) return st;
  }
  {
    StageScope sc(c, ST_UPDATE);
    launch_update_positions(c->stream, c->d_reps, c->d_slot_rep, s0, ns, c->rx, c->ry, c->rz, c->mf, c->bz, c->dxv,
                            c->dyv, c->dzv, c->dsp, c->d_err);
    charge_log_internal(c);
    c->pos_version++;
  }
  HCK(hipGetLastError());
  return 0;
}

int smdlt_step(smdlt_ctx* c, uint64_t t, const double* noise) {
  if (!c) return SMDLT_E_ARG;
  if (!c->have_pos) return SMDLT_E_STATE;
  HCK(hipSetDevice(c->device));
  int st = 0;  // lists are (re)built after the remap, in the wrap frame
  if (noise) {
    st = gather_aos(c, noise, c->tmp);
    if (st) return st;
  }
  st = step_internal(c, t, noise ? c->tmp : nullptr, true);
  if (st) return st;
  st = check_err(c);
  if (st) return st;
  int conv = c->P[c->cur_rep_begin].noise_mode == SMDLT_NOISE_LANCZOS ? read_iters(c, nullptr, nullptr) : 0;
  int g = grow_lists(c);
  collect_timers(c);
  if (g < 0) return g;
  if (g > 0) return SMDLT_E_NLIST_OVERFLOW;
  return conv;
}

// capture the work that body() enqueues on the context stream as a graph
static int capture_graph(smdlt_ctx* c, hipGraphExec_t* out, const std::function<int()>& body) {
  hipGraph_t graph;
  bool prof = c->prof;
  c->prof = false;
  HCK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  int st = body();
  hipError_t e = hipStreamEndCapture(c->stream, &graph);
  c->prof = prof;
  if (st) {
    if (e == hipSuccess) hipGraphDestroy(graph);
    return st;
  }
  HCK(e);
  e = hipGraphInstantiate(out, graph, nullptr, nullptr, 0);
  hipGraphDestroy(graph);
  HCK(e);
  return 0;
}

#define LZ_CHUNK 8

// positions and displacement trackers <-> the saved copy (a graph step whose
// neighbour lists overflowed is undone and redone eagerly with larger lists)
static void save_state(smdlt_ctx* c, bool restore) {
  double* cur[] = {c->rx, c->ry, c->rz, c->dxv, c->dyv, c->dzv, c->dsp};
  long long len[] = {c->nst, c->nst, c->nst, c->nst, c->nst, c->nst, 3LL * c->nst};
  long long off = 0;
  for (int k = 0; k < 7; ++k) {
    if (restore)
      launch_copy_f64(c->stream, cur[k], c->saved + off, len[k]);
    else
      launch_copy_f64(c->stream, c->saved + off, cur[k], len[k]);
    off += len[k];
  }
  if (restore)
    launch_copy_f64(c->stream, c->tau, c->saved + off, 3LL * c->nrep);
  else
    launch_copy_f64(c->stream, c->saved + off, c->tau, 3LL * c->nrep);
  if (c->nq_total > 0) {
    if (restore)
      launch_copy_f64(c->stream, c->qdisp, c->qdisp_saved, 3LL * c->nq_total);
    else
      launch_copy_f64(c->stream, c->qdisp_saved, c->qdisp, 3LL * c->nq_total);
  }
}

// The graphs of one step for the current replica range (see smdlt_ctx).
static int build_step_graphs(smdlt_ctx* c, uint64_t t) {
  int rb = c->cur_rep_begin, re = c->cur_rep_end;
  int s0 = c->rh[rb].slot0, ns = c->rh[re - 1].slot0 + pad_slots(c->rh[re - 1].n) - s0;
  const smdlt_params& p = c->P[rb];
  const double tol = p.lanczos_tol > 0 ? p.lanczos_tol : 1e-6;
  const int first = std::min(c->maxit, LZ_CHUNK);
  int st = capture_graph(c, &c->gexec, [&]() -> int {
    save_state(c, false);
    launch_remap(c->stream, c->d_reps, c->d_slot_rep, s0, ns, c->rx, c->ry, c->rz, c->dxv, c->dyv, c->dzv, c->dsp, 1);
    int s = neighbors_internal(c, 0);
    if (!s) s = forces_internal(c, true);
    if (s) return s;
    charge_rel_internal(c);
    launch_noise(c->stream, c->d_slot_rep, c->d_reps, s0, ns, c->d_seeds, t, c->z);
    mobility_internal(c, c->f, c->mf, nullptr, nullptr, nullptr, nullptr);
    lz_start(c, c->z);
    for (int j = 1; j <= first; ++j) lz_iteration(c, c->z, j, tol, c->maxit);
    return 0;
  });
  if (st) return st;
  for (int j0 = first; j0 < c->maxit; j0 += LZ_CHUNK) {
    hipGraphExec_t g = nullptr;
    st = capture_graph(c, &g, [&]() -> int {
      for (int j = j0 + 1; j <= std::min(c->maxit, j0 + LZ_CHUNK); ++j) lz_iteration(c, c->z, j, tol, c->maxit);
      return 0;
    });
    if (st) return st;
    c->gchunks.push_back(g);
  }
  return capture_graph(c, &c->gtail, [&]() -> int {
    lz_finish(c, c->bz);
    launch_update_positions(c->stream, c->d_reps, c->d_slot_rep, s0, ns, c->rx, c->ry, c->rz, c->mf, c->bz, c->dxv,
                            c->dyv, c->dzv, c->dsp, c->d_err);
    charge_log_internal(c);
    c->pos_version++;
    return 0;
  });
}

int smdlt_run(smdlt_ctx* c, uint64_t t0, int nsteps, int use_graph) {
  if (!c || nsteps < 0) return SMDLT_E_ARG;
  if (!c->have_pos) return SMDLT_E_STATE;
  HCK(hipSetDevice(c->device));
  int st = 0;
  for (int k = 0; k < nsteps; ++k) {
    uint64_t t = t0 + k;
    const bool jump = c->lat_t != UINT64_MAX && t != c->lat_t && t != c->lat_t + 1;
    if (!use_graph || !c->lists_built || c->force_rebuild || jump ||
        c->P[c->cur_rep_begin].noise_mode != SMDLT_NOISE_LANCZOS) {
      // the first step builds the lists after its remap; TEA and free-draining steps run eagerly.
      // An eager step checks for list overflow on the host and grows the lists before its forces.
      // A step after a jump in strain rebuilds its lists eagerly (a captured step never forces it).
      st = step_internal(c, t, nullptr, true);
      if (st) return st;
      continue;
    }
    // lattice for this step, then replay the captured step
    std::vector<HostLattice> lat_before = c->lat;
    std::vector<double> ct_before = c->ct_prev, st_before = c->st_prev;
    st = lattice_to_device(c, t);
    if (st) return st;
    if (!c->gexec) {
      st = build_step_graphs(c, t);
      if (st) {
        destroy_graphs(c);
        return st;
      }
    }
    StageScope sc(c, ST_MOB);
    HCK(hipGraphLaunch(c->gexec, c->stream));
    // chunks the previous step needed go in unchecked; then poll between chunks
    int j0 = std::min(c->maxit, LZ_CHUNK), ch = 0, mx = 0, ov = 0;
    while (ch < (int)c->gchunks.size() && j0 < c->expect_iters) {
      HCK(hipGraphLaunch(c->gchunks[ch++], c->stream));
      j0 += LZ_CHUNK;
    }
    for (;;) {
      int all = lz_poll(c, &mx, &ov);
      if (all < 0) return all;
      if (ov || all || ch >= (int)c->gchunks.size()) break;
      HCK(hipGraphLaunch(c->gchunks[ch++], c->stream));
    }
    if (ov) {  // truncated lists: undo, grow, redo this step eagerly
      save_state(c, true);
      int g = grow_lists(c);
      if (g < 0) return g;
      c->force_rebuild = true;
      c->lat = lat_before;
      c->ct_prev = ct_before;
      c->st_prev = st_before;
      st = step_internal(c, t, nullptr, true);
      if (st) return st;
      continue;
    }
    c->expect_iters = mx;
    HCK(hipGraphLaunch(c->gtail, c->stream));
  }
  HCK(hipStreamSynchronize(c->stream));
  st = check_err(c);
  if (st) return st;
  int conv = c->P[c->cur_rep_begin].noise_mode == SMDLT_NOISE_LANCZOS ? read_iters(c, nullptr, nullptr) : 0;
  int g = grow_lists(c);
  collect_timers(c);
  if (g < 0) return g;
  if (g > 0) return SMDLT_E_NLIST_OVERFLOW;
  return conv;
}

int smdlt_set_profiling(smdlt_ctx* c, int on) {
  if (!c) return SMDLT_E_ARG;
  collect_timers(c);
  c->prof = on != 0;
  for (int s = 0; s < ST_COUNT; ++s) {
    c->stage_ms[s] = 0;
    c->stage_cnt[s] = 0;
  }
  return 0;
}

int smdlt_stage_times(const smdlt_ctx* c, double* ms, int* counts, int max_stages) {
  if (!c || !ms) return SMDLT_E_ARG;
  collect_timers(const_cast<smdlt_ctx*>(c));
  int n = std::min(max_stages, (int)ST_COUNT);
  for (int s = 0; s < n; ++s) {
    ms[s] = c->stage_ms[s];
    if (counts) counts[s] = c->stage_cnt[s];
  }
  return n;
}

const char* smdlt_stage_name(int s) { return (s >= 0 && s < ST_COUNT) ? kStageNames[s] : ""; }

int smdlt_sync(smdlt_ctx* c) {
  if (!c) return SMDLT_E_ARG;
  HCK(hipStreamSynchronize(c->stream));
  return 0;
}

int smdlt_upload_charges(smdlt_ctx* c, const int* nq, const int* track, const int* sign) {
  if (!c || !nq) return SMDLT_E_ARG;
  HCK(hipSetDevice(c->device));
  int tot = 0, mx = 0, maxn = 0;
  for (int r = 0; r < c->nrep; ++r) {
    if (nq[r] < 0 || nq[r] > c->rh[r].n) return SMDLT_E_ARG;
    c->q0[r] = tot;
    c->nq[r] = nq[r];
    tot += nq[r];
    mx = std::max(mx, nq[r]);
    maxn = std::max(maxn, c->rh[r].n);
  }
  for (int k = 0; k < tot; ++k) {
    if (!track) return SMDLT_E_ARG;
  }
  int k = 0;
  for (int r = 0; r < c->nrep; ++r) {
    std::vector<char> seen(c->rh[r].n, 0);
    for (int q = 0; q < nq[r]; ++q, ++k) {
      int b = track[k];
      if (b < 0 || b >= c->rh[r].n || seen[b]) return SMDLT_E_ARG;
      seen[b] = 1;
      if (sign && sign[k] != 1 && sign[k] != -1) return SMDLT_E_ARG;
    }
  }
  void* ptrs[] = {c->qtrack, c->qsign, c->qold, c->occ, c->log_off, c->mh_feats, c->mh_log, c->qdisp, c->qdisp_saved,
                  c->mc_start, c->mc_jf, c->mc_de, c->mc_state};
  for (void* p : ptrs)
    if (p) hipFree(p);
  c->qtrack = c->qsign = c->qold = c->log_off = c->mh_log = c->mc_start = c->mc_jf = nullptr;
  c->mc_de = nullptr;
  c->mc_state = nullptr;
  c->qdisp = c->qdisp_saved = nullptr;
  c->occ = nullptr;
  c->mh_feats = nullptr;
  c->mh_swept = false;
  c->mh_feats_valid = false;
  int st = 0;
  st |= alloc_dev(&c->qtrack, tot);
  st |= alloc_dev(&c->mh_feats, 16 * (size_t)std::max(1, tot));  // MH sweeps' attempt records
  st |= alloc_dev(&c->mh_log, 6 * (size_t)std::max(1, tot));
  st |= alloc_dev(&c->qsign, tot);
  st |= alloc_dev(&c->qold, tot);
  st |= alloc_dev(&c->log_off, c->nrep);
  st |= alloc_dev(&c->qdisp, 3 * (size_t)std::max(1, tot));
  st |= alloc_dev(&c->qdisp_saved, 3 * (size_t)std::max(1, tot));
  c->occ_stride = ((long long)maxn + 255) / 256 * 256;
  st |= alloc_dev(&c->occ, (size_t)c->occ_stride * c->nrep);
  st |= alloc_dev(&c->mc_start, std::max(1, tot));
  st |= alloc_dev(&c->mc_jf, std::max(1, tot));
  st |= alloc_dev(&c->mc_de, 3 * (size_t)std::max(1, tot));
  if (!st && hipMalloc(&c->mc_state, mc_state_bytes() * c->nrep) != hipSuccess) st = 1;
  if (!c->mh_cnt && !st) {  // the MH sweeps' per-bead and per-replica arrays (sized by the context)
    const long long ns1 = (long long)c->nst + 1;
    st |= alloc_dev(&c->mh_cnt, ns1);
    st |= alloc_dev(&c->mh_off, ns1);
    st |= alloc_dev(&c->mh_hcnt, ns1);
    st |= alloc_dev(&c->mh_hoff, ns1);
    st |= alloc_dev(&c->mh_V, (size_t)std::max(1, c->nst));
    st |= alloc_dev(&c->mh_nmoves, c->nrep);
    st |= alloc_dev(&c->mh_done, c->nrep);
    st |= alloc_dev(&c->mh_stop, c->nrep);
    if (!st && hipMalloc(&c->mh_moves, 16 * 64 * (size_t)c->nrep) != hipSuccess) st = 1;
  }
  if (st) return SMDLT_E_NOMEM;
  std::vector<int> sg(tot, 1);
  if (sign) sg.assign(sign, sign + tot);
  if (tot) {
    HCK(hipMemcpy(c->qtrack, track, 4 * tot, hipMemcpyHostToDevice));
    HCK(hipMemcpy(c->qold, track, 4 * tot, hipMemcpyHostToDevice));
    HCK(hipMemcpy(c->qsign, sg.data(), 4 * tot, hipMemcpyHostToDevice));
  }
  HCK(hipMemcpy(c->log_off, c->q0.data(), 4 * c->nrep, hipMemcpyHostToDevice));
  HCK(hipMemset(c->qdisp, 0, sizeof(double) * 3 * (size_t)std::max(1, tot)));
  for (int r = 0; r < c->nrep; ++r) {
    c->rh[r].q0 = c->q0[r];
    c->rh[r].nq = c->nq[r];
  }
  c->nq_total = tot;
  c->max_nq = mx;
  destroy_graphs(c);
  return upload_reps(c);
}

int smdlt_download_charges(smdlt_ctx* c, int* track, int* old) {
  if (!c) return SMDLT_E_ARG;
  HCK(hipSetDevice(c->device));
  if (c->nq_total == 0) return 0;
  if (track) HCK(hipMemcpy(track, c->qtrack, 4 * c->nq_total, hipMemcpyDeviceToHost));
  if (old) HCK(hipMemcpy(old, c->qold, 4 * c->nq_total, hipMemcpyDeviceToHost));
  return 0;
}

// The sweep of smdlt_mc_sweep and smdlt_mc_sweep_noalloc on device buffers:
// du (the replicas' host-fed uniforms, or null), dcons (nrep ints) and dlog
// (reference mode: 6 ints per charge, or null).  The MH modes log into the
// context's own buffers.
// The neighbourhood lists of the MH sweeps for the active replicas at the
// current positions: counts, offsets by an exclusive scan, then the lists.
// may_grow = false (the allocation-free sweep): lists that outgrow the capacity
// an earlier sweep reserved are an error, not a reallocation.
static int mh_build_lists(smdlt_ctx* c, int rb, int re, int max_n, bool may_grow) {
  const long long ns1 = (long long)c->nst + 1;
  if (!c->mh_cnt) return SMDLT_E_STATE;
  launch_fill_i32(c->stream, c->mh_cnt, 0, ns1);
  launch_fill_i32(c->stream, c->mh_hcnt, 0, ns1);
  launch_mh_lists(c->stream, c->d_reps, rb, re - rb, max_n, c->rx, c->ry, c->rz, c->mh_cnt, nullptr, nullptr,
                  c->mh_hcnt, nullptr, nullptr, false);
  launch_scan(c->stream, c->mh_cnt, c->mh_off, (int)ns1);
  launch_scan(c->stream, c->mh_hcnt, c->mh_hoff, (int)ns1);
  int total[2] = {0, 0};
  HCK(hipMemcpyAsync(&total[0], c->mh_off + c->nst, 4, hipMemcpyDeviceToHost, c->stream));
  HCK(hipMemcpyAsync(&total[1], c->mh_hoff + c->nst, 4, hipMemcpyDeviceToHost, c->stream));
  HCK(hipStreamSynchronize(c->stream));
  if (total[0] < 0 || total[1] < 0) return SMDLT_E_GEOMETRY;  // more than 2^31 entries
  int** lists[2] = {&c->mh_list, &c->mh_hlist};
  long long* caps[2] = {&c->mh_list_cap, &c->mh_hlist_cap};
  for (int l = 0; l < 2; ++l) {
    if (total[l] <= *caps[l]) continue;
    if (!may_grow) return SMDLT_E_STATE;
    if (*lists[l]) hipFree(*lists[l]);
    *lists[l] = nullptr;
    *caps[l] = 0;
    const long long cap = (long long)total[l] + total[l] / 4 + 1024;
    if (alloc_dev(lists[l], cap)) return SMDLT_E_NOMEM;
    *caps[l] = cap;
  }
  launch_mh_lists(c->stream, c->d_reps, rb, re - rb, max_n, c->rx, c->ry, c->rz, nullptr, c->mh_off, c->mh_list,
                  nullptr, c->mh_hoff, c->mh_hlist, true);
  return 0;
}

static int mc_sweep_internal(smdlt_ctx* c, int mode, const float* du, int n_uniforms, int* dcons, int* dlog,
                             int* consumed, int* log, bool may_alloc) {
  const bool mh = mode != SMDLT_MC_REFERENCE;
  if (mh) dlog = c->mh_log;
  HCK(hipMemsetAsync(c->mc_err, 0, 4, c->stream));
  int rb = c->cur_rep_begin, re = c->cur_rep_end;
  if (mh) {
    int max_n = 0, max_nq = 0;
    for (int r = rb; r < re; ++r) {
      if (c->nq[r] > 0) max_n = std::max(max_n, c->rh[r].n);
      max_nq = std::max(max_nq, c->nq[r]);
    }
    if (max_nq > 0) {
      int st = mh_build_lists(c, rb, re, max_n, may_alloc);
      if (st) return st;
    }
    launch_mh_sweep(c->stream, c->d_reps, rb, re - rb, max_n, max_nq, c->rx, c->ry, c->rz, c->qtrack, c->qold,
                    c->qsign, (signed char*)c->occ, c->occ_stride, du, du ? n_uniforms : 0, c->d_seeds,
                    c->mc_sweeps++, mode == SMDLT_MC_MH_LEARNED ? c->mlp_w : nullptr, dcons, dlog,
                    c->mh_record ? c->mh_feats : nullptr, c->log_off, c->mc_err, c->qdisp, c->mh_off, c->mh_list, c->mh_hoff, c->mh_hlist,
                    c->mh_V, c->mh_moves, c->mh_nmoves,
                    c->mh_done, c->mh_stop);
  } else
    launch_mc_sweep(c->stream, c->d_reps, rb, re - rb, c->max_nq, c->rx, c->ry, c->rz, c->qtrack, c->qold, c->qsign,
                    c->occ, c->occ_stride, c->mc_start, c->mc_jf, c->mc_de, c->mc_state, du, du ? n_uniforms : 0,
                    c->d_seeds, c->mc_sweeps++, dcons, dlog, c->log_off, c->mc_err, c->qdisp);
  HCK(hipGetLastError());
  if (mh) {
    c->mh_swept = true;
    c->mh_feats_valid = c->mh_record;
  }
  HCK(hipStreamSynchronize(c->stream));
  int e = 0;
  HCK(hipMemcpy(&e, c->mc_err, 4, hipMemcpyDeviceToHost));
  if (consumed) HCK(hipMemcpy(consumed, dcons, 4 * c->nrep, hipMemcpyDeviceToHost));
  if (log && dlog) HCK(hipMemcpy(log, dlog, 24 * (size_t)c->nq_total, hipMemcpyDeviceToHost));
  if (e & 8) return SMDLT_E_GEOMETRY;
  return (e & 2) ? SMDLT_E_MC_STREAM : 0;
}

static int mc_sweep_check(smdlt_ctx* c, int mode, const float* uniforms, int n_uniforms) {
  if (!c) return SMDLT_E_ARG;
  if (!c->have_pos) return SMDLT_E_STATE;
  if (mode != SMDLT_MC_REFERENCE && mode != SMDLT_MC_MH_UNIFORM && mode != SMDLT_MC_MH_LEARNED) return SMDLT_E_ARG;
  if (mode == SMDLT_MC_MH_LEARNED && !c->mlp_w) return SMDLT_E_STATE;
  if (uniforms && n_uniforms <= 0) return SMDLT_E_ARG;
  if (mode != SMDLT_MC_REFERENCE && c->nq_total && !c->mh_feats) return SMDLT_E_STATE;
  return 0;
}

int smdlt_mc_sweep(smdlt_ctx* c, int mode, const float* uniforms, int n_uniforms, int* consumed, int* log) {
  int st = mc_sweep_check(c, mode, uniforms, n_uniforms);
  if (st) return st;
  HCK(hipSetDevice(c->device));
  st = upload_reps(c);
  if (st) return st;
  StageScope sc(c, ST_MC);
  float* du = nullptr;
  int* dcons = nullptr;
  int* dlog = nullptr;
  if (uniforms) {
    HCK(hipMalloc(&du, sizeof(float) * (size_t)n_uniforms * c->nrep));
    HCK(hipMemcpy(du, uniforms, sizeof(float) * (size_t)n_uniforms * c->nrep, hipMemcpyHostToDevice));
  }
  HCK(hipMalloc(&dcons, 4 * c->nrep));
  if (mode == SMDLT_MC_REFERENCE && log && c->nq_total) HCK(hipMalloc(&dlog, 24 * (size_t)c->nq_total));
  st = mc_sweep_internal(c, mode, du, n_uniforms, dcons, dlog, consumed, log, true);
  if (du) hipFree(du);
  hipFree(dcons);
  if (dlog) hipFree(dlog);
  return st;
}

// ------------------------------------------------ allocation-free MC sweep
// smdlt_mc_sweep_noalloc stages the host-fed uniforms, the consumed counts and
// the reference-mode log in `scratch`, one file-scope array per GPU, instead of
// allocating them per call.  Its device address is looked up once per device
// by smdlt_create (with that device current) and kept in g_scratch_ptr; the
// array's name is a host-side stand-in and is never passed to a kernel or a
// copy.  Layout: consumed [nrep] ints, the log [6 nq_total] ints, the uniforms
// [nrep][n_uniforms] floats, each part starting on a 256-byte boundary.
static __device__ __align__(256) unsigned char scratch[SMDLT_MC_NOALLOC_MAX_BYTES];
#define SCRATCH_MAX_DEVICES 64
static std::mutex g_scratch_mu;
static unsigned char* g_scratch_ptr[SCRATCH_MAX_DEVICES];

// Called with `device` current.
static int resolve_scratch(int device) {
  if (device < 0 || device >= SCRATCH_MAX_DEVICES) return SMDLT_E_ARG;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return 0;
  void* p = nullptr;
  HCK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  if (!p) return SMDLT_E_HIP;
  g_scratch_ptr[device] = (unsigned char*)p;
  return 0;
}

static size_t align256(size_t n) { return (n + 255) / 256 * 256; }

int smdlt_mc_sweep_noalloc(smdlt_ctx* c, int mode, const float* uniforms, int n_uniforms, int* consumed, int* log) {
  int st = mc_sweep_check(c, mode, uniforms, n_uniforms);
  if (st) return st;
  const size_t b_cons = align256(4 * (size_t)c->nrep);
  const size_t b_log = (mode == SMDLT_MC_REFERENCE) ? align256(24 * (size_t)c->nq_total) : 0;
  const size_t b_u = uniforms ? 4 * (size_t)n_uniforms * c->nrep : 0;
  if (b_cons + b_log + b_u > (size_t)SMDLT_MC_NOALLOC_MAX_BYTES) return SMDLT_E_ARG;
  HCK(hipSetDevice(c->device));
  unsigned char* base = (c->device >= 0 && c->device < SCRATCH_MAX_DEVICES) ? g_scratch_ptr[c->device] : nullptr;
  if (!base) return SMDLT_E_STATE;
  st = upload_reps(c);
  if (st) return st;
  int* dcons = (int*)base;
  int* dlog = b_log ? (int*)(base + b_cons) : nullptr;
  float* du = uniforms ? (float*)(base + b_cons + b_log) : nullptr;
  // the array keeps the previous call's contents: replicas outside the active
  // range report 0 consumed, and skipped log records read as zeros
  HCK(hipMemsetAsync(dcons, 0, b_cons, c->stream));
  if (dlog) HCK(hipMemsetAsync(dlog, 0, b_log, c->stream));
  if (du) HCK(hipMemcpyAsync(du, uniforms, b_u, hipMemcpyHostToDevice, c->stream));
  return mc_sweep_internal(c, mode, du, n_uniforms, dcons, dlog, consumed, log, false);
}

int smdlt_mc_energy(smdlt_ctx* c, double* energy) {
  if (!c || !energy) return SMDLT_E_ARG;
  if (!c->have_pos) return SMDLT_E_STATE;
  HCK(hipSetDevice(c->device));
  int st = upload_reps(c);
  if (st) return st;
  double* de;
  HCK(hipMalloc(&de, 8 * c->nrep));
  HCK(hipMemset(de, 0, 8 * c->nrep));
  launch_mc_energy(c->stream, c->d_reps, c->cur_rep_begin, c->cur_rep_end - c->cur_rep_begin, c->max_nq, c->rx, c->ry,
                   c->rz, c->qtrack, c->qsign, c->mc_de, de);
  HCK(hipStreamSynchronize(c->stream));
  HCK(hipMemcpy(energy, de, 8 * c->nrep, hipMemcpyDeviceToHost));
  hipFree(de);
  return 0;
}

int smdlt_mlp_num_params(void) { return 16 * 64 + 64 + 64 * 64 + 64 + 64 + 1; }

static float bf16_round(float x) {
  uint32_t u;
  memcpy(&u, &x, 4);
  if ((u & 0x7f800000u) != 0x7f800000u) u += 0x7fffu + ((u >> 16) & 1u);
  u &= 0xffff0000u;
  float r;
  memcpy(&r, &u, 4);
  return r;
}

int smdlt_mlp_set_weights(smdlt_ctx* c, const float* w) {
  if (!c || !w) return SMDLT_E_ARG;
  HCK(hipSetDevice(c->device));
  const int np = smdlt_mlp_num_params();
  if (!c->mlp_w && alloc_dev(&c->mlp_w, np)) return SMDLT_E_NOMEM;
  std::vector<float> r(np);
  for (int i = 0; i < np; ++i) r[i] = bf16_round(w[i]);
  HCK(hipMemcpy(c->mlp_w, r.data(), 4 * np, hipMemcpyHostToDevice));
  return 0;
}

int smdlt_mlp_loss_grad(smdlt_ctx* c, const float* w, const float* x, const float* y, int n, float* loss, float* grad) {
  if (!c || !w || n <= 0 || !loss || !grad || (!x) != (!y)) return SMDLT_E_ARG;
  if (!x && n != c->mlp_n) return SMDLT_E_STATE;  // reuse needs the batch of the previous call
  HCK(hipSetDevice(c->device));
  const int np = smdlt_mlp_num_params();
  const int nblk = (n + MLP_TILE - 1) / MLP_TILE;
  if (n > c->mlp_cap) {
    void* old[] = {c->mlp_x, c->mlp_y, c->mlp_part};
    for (void* p : old)
      if (p) hipFree(p);
    c->mlp_x = c->mlp_y = c->mlp_part = nullptr;
    c->mlp_n = 0;
    if (!x) return SMDLT_E_STATE;
    int cap = std::max(n, 1024);
    int cb = (cap + MLP_TILE - 1) / MLP_TILE;
    if (alloc_dev(&c->mlp_x, 16 * (size_t)cap) || alloc_dev(&c->mlp_y, cap) ||
        alloc_dev(&c->mlp_part, (size_t)(np + 1) * cb))
      return SMDLT_E_NOMEM;
    c->mlp_cap = cap;
  }
  if (!c->mlp_ws && (alloc_dev(&c->mlp_ws, np) || alloc_dev(&c->mlp_out, np + 1))) return SMDLT_E_NOMEM;
  std::vector<float> r(np);
  for (int i = 0; i < np; ++i) r[i] = bf16_round(w[i]);
  HCK(hipMemcpyAsync(c->mlp_ws, r.data(), 4 * np, hipMemcpyHostToDevice, c->stream));
  if (x) {
    HCK(hipMemcpyAsync(c->mlp_x, x, 64 * (size_t)n, hipMemcpyHostToDevice, c->stream));
    HCK(hipMemcpyAsync(c->mlp_y, y, 4 * (size_t)n, hipMemcpyHostToDevice, c->stream));
    c->mlp_n = n;
  }
  launch_mlp_loss_grad(c->stream, c->mlp_ws, c->mlp_x, c->mlp_y, n, c->mlp_part, c->mlp_out);
  HCK(hipGetLastError());
  std::vector<float> out(np + 1);
  HCK(hipMemcpyAsync(out.data(), c->mlp_out, 4 * (size_t)(np + 1), hipMemcpyDeviceToHost, c->stream));
  HCK(hipStreamSynchronize(c->stream));
  memcpy(grad, out.data(), 4 * (size_t)np);
  *loss = out[np];
  (void)nblk;
  return 0;
}

int smdlt_mc_features(smdlt_ctx* c, int replica, float* features, int* from, int* to, int* accepted, int cap,
                      int* n) {
  if (!c || replica < 0 || replica >= c->nrep || !n) return SMDLT_E_ARG;
  if (!c->mh_feats || !c->mh_swept || (features && !c->mh_feats_valid)) return SMDLT_E_STATE;
  HCK(hipSetDevice(c->device));
  const int nq = c->nq[replica], off = c->q0[replica];
  *n = nq;
  if (nq > cap) return SMDLT_E_ARG;
  if (features) HCK(hipMemcpy(features, c->mh_feats + 16 * (size_t)off, 64 * (size_t)nq, hipMemcpyDeviceToHost));
  if (from || to || accepted) {
    std::vector<int> lg(6 * (size_t)nq);
    HCK(hipMemcpy(lg.data(), c->mh_log + 6 * (size_t)off, 24 * (size_t)nq, hipMemcpyDeviceToHost));
    for (int k = 0; k < nq; ++k) {
      if (from) from[k] = lg[6 * k + 1];
      if (to) to[k] = lg[6 * k + 2];
      if (accepted) accepted[k] = lg[6 * k + 4];
    }
  }
  return 0;
}

int smdlt_mc_record_features(smdlt_ctx* c, int on) {
  if (!c) return SMDLT_E_ARG;
  c->mh_record = on != 0;
  return 0;
}

// Select a replica sub-range for the per-replica (plain) path: subsequent calls
// touch only replicas [begin, end).  Not part of the reference interface.
int smdlt_select_replicas(smdlt_ctx* c, int begin, int end) {
  if (!c || begin < 0 || end > c->nrep || begin >= end) return SMDLT_E_ARG;
  set_active(c, begin, end);
  destroy_graphs(c);
  return upload_reps(c);
}

}  // extern "C"

int smdlt_stress(smdlt_ctx* c, double* tau, int reset) {
  if (!c || !tau) return SMDLT_E_ARG;
  HCK(hipSetDevice(c->device));
  HCK(hipMemcpyAsync(tau, c->tau, sizeof(double) * 3 * (size_t)c->nrep, hipMemcpyDeviceToHost, c->stream));
  HCK(hipStreamSynchronize(c->stream));
  if (reset) {  // on the context stream, so the next step's sums start after it
    launch_fill_f64(c->stream, c->tau, 0.0, 3LL * c->nrep);
    HCK(hipGetLastError());
    HCK(hipStreamSynchronize(c->stream));
  }
  return 0;
}

int smdlt_chain_observables(smdlt_ctx* c, double* ext, double* com, double* rg) {
  if (!c) return SMDLT_E_ARG;
  if (!c->have_pos) return SMDLT_E_STATE;
  HCK(hipSetDevice(c->device));
  const int rb = c->cur_rep_begin, re = c->cur_rep_end, nk = c->nchains;
  double* d = c->obs;
  launch_chain_obs(c->stream, c->d_reps, rb, re - rb, c->max_chains, c->d_chain0, c->rx, c->ry, c->rz, d, d + nk,
                   d + 4 * (size_t)nk);
  HCK(hipGetLastError());
  std::vector<double> h(10 * (size_t)nk);
  HCK(hipMemcpyAsync(h.data(), d, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c->stream));
  HCK(hipStreamSynchronize(c->stream));
  for (int r = rb; r < re; ++r)
    for (int k = c->chain0[r]; k < c->chain0[r + 1]; ++k) {
      if (ext) ext[k] = h[k];
      for (int q = 0; q < 3; ++q)
        if (com) com[3 * k + q] = h[nk + 3 * k + q];
      for (int q = 0; q < 6; ++q)
        if (rg) rg[6 * k + q] = h[4 * (size_t)nk + 6 * k + q];
    }
  return 0;
}

int smdlt_charge_displacements(smdlt_ctx* c, double* disp, int reset) {
  if (!c || !disp) return SMDLT_E_ARG;
  HCK(hipSetDevice(c->device));
  if (c->nq_total == 0) return 0;
  HCK(hipMemcpyAsync(disp, c->qdisp, sizeof(double) * 3 * (size_t)c->nq_total, hipMemcpyDeviceToHost, c->stream));
  HCK(hipStreamSynchronize(c->stream));
  if (reset) {
    launch_fill_f64(c->stream, c->qdisp, 0.0, 3LL * c->nq_total);
    HCK(hipGetLastError());
    HCK(hipStreamSynchronize(c->stream));
  }
  return 0;
}
