// Shared definitions of the gfx950 timestep library: replica descriptor,
// device helpers and the launch wrappers each kernel file exports to api.cpp.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define SLOT_ALIGN 64       // beads of a replica are padded to whole wavefronts
#define KMAX 512            // largest FFT length (lengths are 2^a 3^b 5^c, multiples of 8)
// complex elements per component per FFT workgroup, by pass; at most FFT_MAXL lines
#ifndef FFT_TILE_X
#define FFT_TILE_X 2048
#endif
#define FFT_TILE_X64 1024  // FP64 grids keep the smaller x tile, as they do in y
#ifndef FFT_TILE_Y
#define FFT_TILE_Y 2048
#endif
#define FFT_TILE_Y64 1024  // FP64 grids: twice the bytes per value, so the 2048 tile would cut occupancy to 3 waves/SIMD
// the z pass has three tiles, the largest that keeps enough workgroups on the
// context's grids: FFT_TILE_Z_LARGE with at least FFT_Z_LARGE_MIN_TASKS of them
// (FP32 grids only: in FP64 it runs at 2 waves/SIMD), FFT_TILE_Z with at least
// FFT_Z_MIN_TASKS, else FFT_TILE_Z_SMALL
#define FFT_TILE_Z_SMALL 256
#define FFT_TILE_Z 512
#define FFT_TILE_Z_LARGE 1024
#define FFT_Z_LARGE_MIN_TASKS 1024
#define FFT_Z_MIN_TASKS 256
#define FFT_MAXL 64
// fused x-y pass: one workgroup of XY_THREADS per (replica, component, z plane)
// when the plane's half spectrum K0*K1/2 fits XY_ELEMS complex values in LDS
#define XY_THREADS 1024
#ifndef XY_ELEMS
#define XY_ELEMS 8192
#endif
#define XY_MAX_POINTS 8.0e6  // fused only while the context's grids total at most this many points
// The five FFT passes run replica chunk by chunk (x, y, z, y, x for one chunk,
// then the next) with chunks of at most this many grid bytes, so a chunk's grid
// stays in the 256 MB last-level cache between its passes
#ifndef FFT_CHUNK_BYTES
#define FFT_CHUNK_BYTES (256.0 * 1048576.0)
#endif
#define FFT_THREADS 256
#define MAX_ORDER 8         // largest B-spline order
#define LZ_MAXIT 128        // largest Lanczos iteration count supported

// Per-replica descriptor, uploaded to the device every step (lattice fields
// change with t).  All lattice quantities are computed on the host in FP64 with
// the reference's expressions so that the FP64 remap kernels reproduce them.
struct RepDev {
  int slot0, n, nb, nc;        // first slot, beads, beads per chain, chains
  int wave0, nwaves;           // wavefront range of the replica's slots
  int flow, pbc_compat, lj_ext, spring_fene;
  int K[3], order;             // PME grid along a1, a2, a3 and B-spline order
  int ncell[3], cell0;         // neighbour-build cell grid and offset
  int q0, nq;                  // charges: offset into the charge arrays, count
  long long grid0;             // element offset of the replica's grid (3 components)
  int bsp0, active;            // B-spline modulus offset; replica enabled for this launch
  int brick0, nbrick[3];       // 8x8x8-point grid bricks of the spreading gather
  int tab0, ntab;              // mixed precision: real-space coefficient table (float4 pairs) and its intervals
  int twoff[3];                // offsets of the twiddle tables exp(-2 pi i t / K[d]) (complex elements)
  int cl0, ncl;                // bead clusters (<= 8 consecutive beads of one chain): first id, count
  // lattice in the wrap frame
  double L, invL, L1xp, L2xp, L2yp, delx;
  double ct, st;               // rotation lab -> wrap frame at this step (KRBC)
  double ct0, st0;             // rotation of the previous step (KRBC round trip)
  double rb[4];                // Lagrange-reduced in-plane basis b1=(rb0,rb1), b2=(rb2,rb3)
  double rd[4];                // its dual rows
  // Ewald split
  double alpha, rc2, rlist2, selfconst, vol;
  double hinv[3];              // s2=y*hinv1, s1=(x-L2xp*s2)*hinv0, s3=z*hinv2
  double kb[4];                // reciprocal basis: b1=(kb0,kb1,0), b2=(0,kb2,0), b3=(0,0,kb3)
  // physics
  double dt, edot, kappas, qmax, kappab, epsilon, lambda_b, kappa_d, r2cut, rv, sqrt2dt;
  double barrier, barrier2;
  double hop2, nbhd2, feat_r, pfloor;  // MH hops: hop radius^2, neighbourhood radius^2, feature radius, floor
  double rnew, skin;           // Verlet trigger (rv-rc)/2 and mobility-list skin
  double halfw;                // half the shortest nonzero lattice vector bound: min(reduced widths, L) / 2
  int treal, reset;            // t % tp and whether this step starts a lattice period
  unsigned long long step;     // step counter for the device noise stream
};

template <typename T> struct cplx_t;
template <> struct cplx_t<float> { typedef float2 type; };
template <> struct cplx_t<double> { typedef double2 type; };

__device__ __forceinline__ float2 cmul(float2 a, float2 b) { return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ __forceinline__ double2 cmul(double2 a, double2 b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
template <typename C> __device__ __forceinline__ C mk2(double x, double y);
template <> __device__ __forceinline__ float2 mk2<float2>(double x, double y) { return make_float2((float)x, (float)y); }
template <> __device__ __forceinline__ double2 mk2<double2>(double x, double y) { return make_double2(x, y); }

// Compensated (TwoSum) accumulator for per-bead velocity sums.
// n / d for n >= 0 with n * d < 2^32 by a multiply-high: with m = floor((2^32 - 1)
// / d) + 1 = ceil(2^32 / d) (2^32 / d when d is a power of two),
// floor(n m / 2^32) = floor(n / d) because n (m d - 2^32) < n d < 2^32.  Built
// once from a wave-uniform divisor, it replaces the per-lane division sequence.
struct FDiv {
  unsigned m;
  int d;
  __device__ __forceinline__ explicit FDiv(int d_) : m(0xFFFFFFFFu / (unsigned)d_ + 1u), d(d_) {}
  __device__ __forceinline__ int div(int n) const { return d == 1 ? n : (int)__umulhi((unsigned)n, m); }
};

// chargeHop's displacement of an accepted hop (S/main26.c:2221-2233): new bead
// minus old bead, reduced by the cell without the shear shift of getNID.
__device__ __forceinline__ void hop_vector(const RepDev& R, const double* rx, const double* ry, const double* rz,
                                           int snew, int sold, double d[3]) {
#pragma clang fp contract(off)
  double dx = rx[snew] - rx[sold], dy = ry[snew] - ry[sold], dz = rz[snew] - rz[sold];
  dx -= R.L2xp * round(dy / R.L2yp);
  dy -= R.L2yp * round(dy / R.L2yp);
  dx -= R.L1xp * round(dx / R.L1xp);
  dz -= R.L * round(dz / R.L);
  d[0] = dx;
  d[1] = dy;
  d[2] = dz;
}

template <typename T> struct CompSum {
  T s, c;
  __device__ __forceinline__ CompSum() : s(0), c(0) {}
  __device__ __forceinline__ void add(T x) {
    T t = s + x;
    T bp = t - s;
    c += (s - (t - bp)) + (x - bp);
    s = t;
  }
  __device__ __forceinline__ double value() const { return (double)s + (double)c; }
};

// Minimum image in the Lagrange-reduced basis (FP64).  For |d| below half the
// reduced cell's smallest width the rounding returns the unique minimum image.
__device__ __forceinline__ void min_image_reduced(const RepDev& R, double& dx, double& dy, double& dz) {
  double f1 = R.rd[0] * dx + R.rd[1] * dy;
  double f2 = R.rd[2] * dx + R.rd[3] * dy;
  double n1 = rint(f1), n2 = rint(f2);
  dx -= n1 * R.rb[0] + n2 * R.rb[2];
  dy -= n1 * R.rb[1] + n2 * R.rb[3];
  dz -= R.L * rint(dz * R.invL);
}

// getNID of the reference, FP64, written in its operation order
// (S/main26.c:1429-1457 shear with the delx shift; X/main14.c:1462-1480 KRBC).
// Returns the pair vector r_i - r_j in the frame the reference returns it in.
__device__ __forceinline__ void ref_nid(const RepDev& R, double xi, double yi, double zi, double xj,
                                        double yj, double zj, double& ox, double& oy, double& oz,
                                        bool to_lab = true) {
#pragma clang fp contract(off)
  double dx = xi - xj, dy = yi - yj, dz = zi - zj;
  double L2xp = R.L2xp, L2yp = R.L2yp;
  if (R.flow == 2 || !R.pbc_compat) {
    dx -= L2xp * round(dy / L2yp);
    dy -= L2yp * round(dy / L2yp);
  } else {
    dx -= L2xp * round(dy / L2yp);
    if (dy > L2yp / 2) {
      dx -= R.delx;
      dy -= L2yp * round(dy / L2yp);
    } else if (dy < -L2yp / 2) {
      dx += R.delx;
      dy -= L2yp * round(dy / L2yp);
    }
  }
  dx -= R.L1xp * round(dx / R.L1xp);
  dz -= R.L * round(dz / R.L);
  if (R.flow == 2 && to_lab) {
    ox = dx * R.ct - dy * R.st;
    oy = dx * R.st + dy * R.ct;
  } else {
    ox = dx;
    oy = dy;
  }
  oz = dz;
}

// Cardinal B-spline weights W[j] = M_P(x + j), j = 0..P-1, x in [0,1); weight j
// belongs to grid point floor(u) - j.  Recursion M_k(y) = [y M_{k-1}(y) +
// (k-y) M_{k-1}(y-1)]/(k-1), unrolled so every index is static.
template <typename T, int P> __device__ __forceinline__ void bspline_weights(T x, T* W) {
  T A[P];
#pragma unroll
  for (int j = 0; j < P; ++j) A[j] = (j == 0) ? (T)1 : (T)0;
#pragma unroll
  for (int k = 2; k <= P; ++k) {
    T div = (T)1 / (T)(k - 1);
#pragma unroll
    for (int j = P - 1; j >= 0; --j) {
      T lo = (j >= 1) ? A[j - 1] : (T)0;
      // one explicit fma, so the rounding does not depend on how the compiler contracts the expression
      A[j] = div * fma(x + (T)j, A[j], ((T)k - x - (T)j) * lo);
    }
  }
#pragma unroll
  for (int j = 0; j < P; ++j) W[j] = A[j];
}

// Beenakker real-space coefficients with the exact overlap correction for r<2
// (a=1, unit self-mobility): M = t1 I + t2 d d^T / r^2.
template <typename T> __device__ __forceinline__ void rpy_real(T r, T alpha, T& t1, T& t2) {
  T rr = r * r, ir = (T)1 / r, ir3 = ir / rr;
  T a2 = alpha * alpha, a3 = a2 * alpha, a5 = a3 * a2, a7 = a5 * a2;
  T ec = erfc(alpha * r);
  T g = exp(-a2 * rr) * (T)0.56418958354775628695;  // 1/sqrt(pi)
  T C1 = (T)0.75 * ir + (T)0.5 * ir3;
  T C2 = (T)4 * a7 * rr * rr + (T)3 * a3 * rr - (T)20 * a5 * rr - (T)4.5 * alpha + (T)14 * a3 + alpha / rr;
  T C3 = (T)0.75 * ir - (T)1.5 * ir3;
  T C4 = -(T)4 * a7 * rr * rr - (T)3 * a3 * rr + (T)16 * a5 * rr + (T)1.5 * alpha - (T)2 * a3 - (T)3 * alpha / rr;
  t1 = C1 * ec + C2 * g;
  t2 = C3 * ec + C4 * g;
  if (r < (T)2) {
    t1 += ((T)1 - (T)0.28125 * r) - C1;
    t2 += (T)0.09375 * r - C3;
  }
}

__device__ __forceinline__ double warp_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ float warp_sum(float v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ double warp_max(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11), counter-based.
__device__ __forceinline__ void philox10(uint32_t c[4], uint32_t k0, uint32_t k1) {
  for (int r = 0; r < 10; ++r) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    uint32_t h0 = (uint32_t)(p0 >> 32), l0 = (uint32_t)p0;
    uint32_t h1 = (uint32_t)(p1 >> 32), l1 = (uint32_t)p1;
    uint32_t n0 = h1 ^ c[1] ^ k0, n2 = h0 ^ c[3] ^ k1;
    c[0] = n0;
    c[1] = l1;
    c[2] = n2;
    c[3] = l0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}
__device__ __forceinline__ double u01_53(uint32_t a, uint32_t b) {
  uint64_t x = ((uint64_t)a << 21) ^ (uint64_t)b;
  x &= ((1ull << 53) - 1);
  return ((double)x + 0.5) * (1.0 / 9007199254740992.0);
}

// --------------------------------------------------------------------------
// Launch wrappers (host) exported by the kernel files.
struct FftTask { int rep, comp, l0, nl; };
struct BrickTask { int rep, b0, b1, b2; };
#define BRICK 8
#define BW_STRIDE (3 * MAX_ORDER)  // B-spline weights stored per bead

struct DevState;  // defined in api.cpp

// k_remap.hip
void launch_remap(hipStream_t s, const RepDev* reps, const int* slot_rep, int slot_begin, int nslots,
                  double* rx, double* ry, double* rz, double* dxv, double* dyv, double* dzv, double* dsp,
                  int mode);
void launch_min_image(hipStream_t s, const RepDev* rep, const double* rx, const double* ry, const double* rz,
                      int npairs, const int* pi, const int* pj, double* out);
void launch_bin_index(hipStream_t s, const RepDev* rep, const double* rx, const double* ry, const double* rz,
                      int npairs, const int* pi, const int* pj, const double* bins, const int* nbins, int* out);
void launch_update_positions(hipStream_t s, const RepDev* reps, const int* slot_rep, int slot_begin, int nslots,
                             double* rx, double* ry, double* rz, const double* mf, const double* bz,
                             double* dxv, double* dyv, double* dzv, double* dsp, int* err);
void launch_disp_max(hipStream_t s, const RepDev* reps, int rep_begin, int nrep, const double* dxv,
                     const double* dyv, const double* dzv, const double* dsp, double* out2);
void launch_noise(hipStream_t s, const int* slot_rep, const RepDev* reps, int slot_begin, int nslots,
                  const uint64_t* seeds, uint64_t step, double* z);
void launch_rotate_vec(hipStream_t s, const RepDev* reps, const int* slot_rep, int slot_begin, int nslots,
                       double* v, int inverse);

// k_nlist.hip
void launch_cell_count(hipStream_t s, const RepDev* reps, const int* slot_rep, int slot_begin, int nslots,
                       const double* rx, const double* ry, const double* rz, int* cell_of, int* cell_cnt,
                       const int* rep_flags);
void launch_scan(hipStream_t s, const int* in, int* out, int n);
void launch_reset_trackers(hipStream_t s, const RepDev* reps, const int* slot_rep, int slot_begin, int nslots,
                           const int* rep_flags, double* dxv, double* dyv, double* dzv, double* dsp);
void launch_nl_check(hipStream_t s, const RepDev* reps, int rep_begin, int nrep, const double* out2, double* aff,
                     int force, int* rep_flags);
void launch_cell_fill(hipStream_t s, const RepDev* reps, const int* slot_rep, int slot_begin, int nslots,
                      const int* cell_of, const int* cell_start, int* cell_fill, int* cell_beads,
                      const int* rep_flags);
void launch_cell_sort(hipStream_t s, const int* cell_start, const int* cell_cnt, int cell_begin, int ncells,
                      const int* unsorted, int* cell_beads);
void launch_build_lists(hipStream_t s, const RepDev* reps, const int* slot_rep, int slot_begin, int nslots,
                        const double* rx, const double* ry, const double* rz, const int* cell_start,
                        const int* cell_cnt, const int* cell_beads, const int* rep_flags, int nslots_total,
                        int cap_rpy, int* nrpy, int* lrpy, int cap_lj, int* nlj, int* llj, int* overflow);

void launch_fill_i32(hipStream_t s, int* p, int v, long long n);
void launch_fill_f64(hipStream_t s, double* p, double v, long long n);
void launch_copy_f64(hipStream_t s, double* dst, const double* src, long long n);
void launch_fill_f32(hipStream_t s, float* p, float v, long long n);

// k_forces.hip
void launch_forces(hipStream_t s, const RepDev* reps, const int* slot_rep, int slot_begin, int nslots,
                   const double* rx, const double* ry, const double* rz, int nslots_total, const int* nlj,
                   const int* llj, double* f, double* vir, int* err);
void launch_charge_rel(hipStream_t st, const RepDev* reps, int rep_begin, int nrep, int max_chains, const double* rx,
                       const double* ry, const double* rz, double* rel);
void launch_charge_log(hipStream_t st, const RepDev* reps, int rep_begin, int nrep, int max_nq, const int* qtrack,
                       const double* mf, const double* bz, const double* rel, double* qdisp);
void launch_virial_acc(hipStream_t st, const RepDev* reps, int rep_begin, int nrep, const double* vir, double* tau);
void launch_chain_obs(hipStream_t st, const RepDev* reps, int rep_begin, int nrep, int max_chains, const int* chain0,
                      const double* rx, const double* ry, const double* rz, double* ext, double* com, double* rg);
void launch_coulomb(hipStream_t s, const RepDev* reps, int rep_begin, int nrep, const double* rx,
                    const double* ry, const double* rz, const int* qtrack, const int* qsign, double* f, int max_nq);

// k_mobility.hip
// Workgroups are dealt to the 8 XCDs round-robin by workgroup id.  This maps
// the ids so that XCD x receives one contiguous range of logical blocks (its
// beads' neighbours then live in its own L2).
__device__ __forceinline__ int xcd_block(int b, int nb) {
  const int x = b & 7, i = b >> 3, q = nb >> 3, rem = nb & 7;
  return x * q + (x < rem ? x : rem) + i;
}

struct MobArgs {
  const RepDev* reps;
  const int* slot_rep;
  int slot_begin, nslots, nst;
  const double *rx, *ry, *rz;
  const int *nrpy, *lrpy;   // mobility list, row-major: lrpy[slot * cap_rpy + k]
  int cap_rpy;
  const double* v;          // input vector (AoS 3N); scaled by 1/beta[r] when beta != nullptr
  const double* beta;       // per-replica divisor (Lanczos normalisation) or nullptr
  double* vout;             // where the (scaled) input is stored for the real-space pass
  double* u;                // output M v (AoS 3N)
  double* adot;             // cluster path: per-cluster partial v.u of the product, the Lanczos alpha (nullptr: skip)
  const int* done;          // per-replica skip flag (nullptr: none)
  void* grid;               // grid buffer (T)
  long long grid_begin, grid_len;
  const void* bsp;          // B-spline moduli (T)
  const void* tw;           // twiddles (complex T, KMAX)
  const FftTask *tx, *ty, *tz, *txy;
  int ntx, nty, ntz, ntxy;
  int ztile;                // complex elements per component of a z task: FFT_TILE_Z_SMALL, FFT_TILE_Z or FFT_TILE_Z_LARGE
  int order;                // B-spline order (4, 6 or 8), one per context
  hipEvent_t* ev;           // 9 events bracketing the 8 launches (profiling) or nullptr
  // spreading gather: per-bead stencil base and weights, beads binned by brick
  int* bbase;               // 3 ints per slot
  void* bw;                 // BW_STRIDE weights (T) per slot
  int *brick_cnt, *brick_start, *brick_fill, *brick_beads;
  const BrickTask* bricks;
  int nbricks, brick_begin, brick_len;
  int prep;                 // recompute base/weights/binning (positions or lattice changed)
  int* scan_tmp;
  // mixed precision: FP32 copies of the positions (per position version) and of
  // the scaled input vector (per product), 16-byte aligned for one load per bead
  float4* p32;
  float4* v32;
  // mixed precision: per brick, its candidate beads for the matrix-core spread
  // (bcand[bcand_start[b] ..] , bcand_cnt[b] entries; rebuilt with the bricks)
  int *bcand_cnt, *bcand_start, *bcand;
  // mixed precision: per replica, for r^2 in [RTAB_R0^2, rc^2) on the segments of
  // rtab_locate, cubic coefficients (in its s) of t1(r) and of t2(r)/r^2 with the
  // RPY overlap correction below r = 2 (r^2 = 4 is a knot): rtab[2i], rtab[2i+1];
  // the last entry is zero
  const float4* rtab;
  int phase;                // bit 0: bead preparation and input scaling; bit 1: spread .. final
  int scale_in_spread;      // the spread stages v / beta itself and stores vout, v32 (set by the launcher)
  // mixed precision real space by cluster pairs: cluster c holds beads cl_first[c] ..
  // + cl_cnt[c] - 1; lcl[c * capc + k], k < nclp[c]: clusters with a bead pair in the
  // list, as (first slot << 4) | count
  const int *cl_first, *cl_cnt, *lcl, *nclp;
  int capc, cl_begin, ncl;
  int ntab_max;             // largest coefficient table of the range (entries of 2 float4)
  // host memory, read by the launcher only: FFT task offsets (x, y, z, x-y) of the
  // nfch + 1 chunk boundaries, relative to tx, ty, tz, txy
  const int* fch;
  int nfch;
  hipEvent_t* evc;          // profiling with several chunks: 6 events per chunk around its 5 passes
};
#define RTAB_LDS_MAX 1280   // tables up to this many entries (40 KB) are staged in LDS
#define CL_LIST_PAD 8       // entries past the last cluster-pair row, so a row's 4-entry groups read in bounds
#define SCALE_FUSE_MAX_SLOTS 32768  // ranges up to this many slots normalise the input in the spread
#define RTAB_R0 1.0f  // pairs closer than this take the FP32 closed form
// Real-space table segments: r^2 in [1, rc^2) cut into 2^RTAB_MBITS equal
// segments per octave, so the segment is the float's exponent and top RTAB_MBITS
// mantissa bits and the position in it is the remaining mantissa under exponent 0
// (s in [1, 2), linear in r^2).  The spacing in r is r / 2^(RTAB_MBITS + 1): fine
// near contact where the coefficients vary fastest, and the lookup needs no square
// root or float conversion.  With RTAB_MBITS 5 the FP32 cubic is within ~1e-7 of
// the coefficients' largest value for ewald_alpha 0.1 .. 0.9 and rc 5 .. 48.
#define RTAB_MBITS 5
#define RTAB_SHIFT (23 - RTAB_MBITS)
#define RTAB_BASE (0x3f800000u >> RTAB_SHIFT)  // the segment starting at r^2 = RTAB_R0^2 = 1
// A replica's table holds ntab - 1 segments from r^2 = 1 up to the one holding
// rc^2, then one entry of zero coefficients.  rtab_top(ntab) is the first r^2 of
// that zero entry; r^2 clamped to [1, rtab_top] selects an entry (the zero one for
// pairs that are not summed) and its position s.
__host__ __device__ __forceinline__ float rtab_top(int ntab) {
  return __builtin_bit_cast(float, (unsigned)(RTAB_BASE + (unsigned)(ntab - 1)) << RTAB_SHIFT);
}
__device__ __forceinline__ void rtab_locate(float rrc, int& it, float& s) {
  const unsigned b = __float_as_uint(rrc);
  it = (int)((b >> RTAB_SHIFT) - RTAB_BASE);
  s = __uint_as_float(((b << RTAB_MBITS) & 0x007fffffu) | 0x3f800000u);
}

// Overlap-corrected real-space coefficients in FP32 for r < RTAB_R0, written with
// erf so that the 1/r^3 terms of the Ewald and RPY far forms cancel analytically:
// t1 = (1 - 9r/32) - C1 erf(alpha r) + C2 g, t2 = 3r/32 - C3 erf(alpha r) + C4 g.
__device__ __forceinline__ void rpy_real_near_f32(float r, float rr, float alpha, float& t1, float& t2) {
  const float ir = 1.f / r, ir3 = ir / rr;
  const float a2 = alpha * alpha, a3 = a2 * alpha, a5 = a3 * a2, a7 = a5 * a2;
  const float ef = erff(alpha * r), g = expf(-a2 * rr) * 0.56418958354775628695f;
  const float C1 = 0.75f * ir + 0.5f * ir3, C3 = 0.75f * ir - 1.5f * ir3;
  const float C2 = 4.f * a7 * rr * rr + 3.f * a3 * rr - 20.f * a5 * rr - 4.5f * alpha + 14.f * a3 + alpha / rr;
  const float C4 = -4.f * a7 * rr * rr - 3.f * a3 * rr + 16.f * a5 * rr + 1.5f * alpha - 2.f * a3 - 3.f * alpha / rr;
  t1 = (1.f - 0.28125f * r) - C1 * ef + C2 * g;
  t2 = 0.09375f * r - C3 * ef + C4 * g;
}
template <typename T> void launch_mobility(hipStream_t s, const MobArgs& a);

// One replica's parameters for the TEA sample (ewaldBin): lattice vectors as the
// reference holds them, the Ewald split it uses (M_HI = 3.5) and its k table range.
#define TEA_THREADS 256
#define TEA_MAXK 512
struct TeaRep {
  int slot0, n, bead0, k0, nk, nmax;
  double L1[2], L2[2], L, rce, alpha, alpha2, alpha3, alpha5, alpha7;
  double dtr[3];
};
void launch_tea_rows(hipStream_t s, const TeaRep* reps, const int* bead_rep, int nbeads, const double* rx,
                     const double* ry, const double* rz, const double* ktab, double* eps_out, double* rowsum_out);
void launch_far_table(hipStream_t s, const TeaRep* phases, int nphase, const double* ktab, const int nb[3],
                      const double bs[3], double* table);

void launch_build_clusters(hipStream_t st, const RepDev* reps, const int* slot_rep, int cl_begin, int ncl,
                           int max_words, const int* cl_first, const int* cl_cnt, const int* slot_cl,
                           const int* rep_flags, const int* nrpy, const int* lrpy, int cap_rpy, int capc, int* nclp,
                           int* lcl, int* overflow);

// k_lanczos.hip
void lz_wave_dot(hipStream_t st, const RepDev* reps, const int* slot_rep, int slot_begin, int nslots, const double* x,
                 const double* y, const int* done, double* partial);
void lz_rep_reduce(hipStream_t st, const RepDev* reps, int rep_begin, int nrep, const double* partial,
                   const int* done, double* out, int stride, int off, int take_sqrt, double* copy_to);
void lz_orth(hipStream_t st, const RepDev* reps, const int* slot_rep, int slot_begin, int nslots, double* w,
             const double* vj, const double* vjm1, double* scal, int j, const int* done, const double* partial_alpha,
             const double* cluster_alpha, double* partial);
void lz_tri(hipStream_t st, const RepDev* reps, int rep_begin, int nrep, const double* partial, double* scal, int j,
            double tol, int maxit, int* done, double* bdiv, int* iters);
void lz_combine(hipStream_t st, const RepDev* reps, const int* slot_rep, int slot_begin, int nslots, const double* V,
                int nst, const double* scal, const int* iters, double* y);
void lz_tea(hipStream_t st, const int* slot_rep, int slot_begin, int nslots, const double* mz, const double* z,
            const double* diag, const double* beta, const double* C, double* y);
#define SCAL_STRIDE (4 * LZ_MAXIT + 8)
void lz_reset_checks(hipStream_t st, double* scal, int rep_begin, int nrep);

// k_mc.hip
#define MLP_TILE 64
void launch_mlp_loss_grad(hipStream_t st, const float* w, const float* X, const float* Y, int n, float* part,
                          float* out);
// MH sweeps: per-bead neighbourhood and hop lists (count, then fill at the
// scanned offsets), then the sweep in rounds; V (nst doubles), moves (nrep x 64 int4),
// nmoves/done/stop (nrep ints) are its working arrays
void launch_mh_lists(hipStream_t st, const RepDev* reps, int rep_begin, int nrep, int max_n, const double* rx,
                     const double* ry, const double* rz, int* cnt, const int* off, int* list, int* hcnt,
                     const int* hoff, int* hlist, bool fill);
void launch_mh_sweep(hipStream_t st, const RepDev* reps, int rep_begin, int nrep, int max_n, int max_nq,
                     const double* rx, const double* ry, const double* rz, int* track, int* old, const int* sign,
                     signed char* occ, long long occ_stride, const float* uniforms, int nu, const uint64_t* seeds,
                     uint64_t sweep_id, const float* weights, int* consumed, int* log, float* feats,
                     const int* log_off, int* err, double* qdisp, const int* nb_off, const int* nb_list,
                     const int* hop_off, const int* hop_list, double* V, void* moves, int* nmoves, int* done,
                     int* stop);
// reference-mode sweep; start/jfirst (nq_total ints), de0 (3 nq_total doubles)
// and state (nrep x mc_state_bytes()) are the sweep's working arrays
size_t mc_state_bytes();
void launch_mc_sweep(hipStream_t st, const RepDev* reps, int rep_begin, int nrep, int max_nq, const double* rx,
                     const double* ry, const double* rz, int* track, int* old, const int* sign, unsigned char* occ,
                     long long occ_stride, int* start, int* jfirst, double* de0, void* state, const float* uniforms,
                     int nu, const uint64_t* seeds, uint64_t sweep_id, int* consumed, int* log, const int* log_off,
                     int* err, double* qdisp);
// rows: nq_total doubles of working space
void launch_mc_energy(hipStream_t st, const RepDev* reps, int rep_begin, int nrep, int max_nq, const double* rx,
                      const double* ry, const double* rz, const int* track, const int* sign, double* rows,
                      double* e);
