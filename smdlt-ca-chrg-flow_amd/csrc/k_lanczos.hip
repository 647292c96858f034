// Lanczos (Krylov) Brownian displacement y = M^(1/2) z per replica, FP64:
//   v_1 = z/|z|;  w = M v_j;  a_j = v_j.w;  w -= a_j v_j + b_{j-1} v_{j-1};  b_j = |w|
//   c_j = |z| T_j^(1/2) e_1 (T_j tridiagonal), stop when |c_j - [c_{j-1};0]| < tol |c_j|,
//   y = V_j c_j.
// This replaces the TEA product of updateHI (S/main26.c:787-804).  Reductions
// are two-level and ordered (per-wavefront partials, then one wavefront per
// replica), so results do not depend on scheduling.  The tridiagonal square
// root is evaluated by a quadrature of the inverse-square-root integral on one
// workgroup per replica (k_lz_tri).
#include "common.h"

#define SCAL_STRIDE (4 * LZ_MAXIT + 8)
#define S_AL 0
#define S_BE LZ_MAXIT
#define S_C (2 * LZ_MAXIT)
#define S_CP (3 * LZ_MAXIT)
#define S_B0 (4 * LZ_MAXIT)
#define S_IT (4 * LZ_MAXIT + 1)
#define S_ERR (4 * LZ_MAXIT + 2)
#define S_JMIN (4 * LZ_MAXIT + 4)  // first iteration whose square root is evaluated

// per-wave partial of x.y (y == nullptr: x.x)
__global__ void k_wave_dot(const RepDev* __restrict__ reps, const int* __restrict__ slot_rep, int slot_begin,
                           int nslots, const double* x, const double* y, const int* done, double* partial) {
  int s = slot_begin + blockIdx.x * blockDim.x + threadIdx.x;
  bool in = s < slot_begin + nslots;
  int r = in ? slot_rep[s] : -1;
  double v = 0.0;
  bool act = false;
  if (r >= 0) {
    const RepDev& R = reps[r];
    act = R.active && !(done && done[r]);
    if (act && s - R.slot0 < R.n) {
      for (int c = 0; c < 3; ++c) v += x[3 * s + c] * (y ? y[3 * s + c] : x[3 * s + c]);
    }
  }
  v = warp_sum(v);
  if ((threadIdx.x & 63) == 0 && act) partial[s >> 6] = v;
}

// one wavefront per replica: sum of its partials in wave order -> out[r*stride + off]
__global__ void k_rep_reduce(const RepDev* __restrict__ reps, int rep_begin, const double* partial, const int* done,
                             double* out, int stride, int off, int take_sqrt, double* copy_to) {
  int r = rep_begin + blockIdx.x;
  const RepDev& R = reps[r];
  if (!R.active || (done && done[r])) return;
  double v = 0.0;
  for (int w = threadIdx.x; w < R.nwaves; w += 64) v += partial[R.wave0 + w];
  v = warp_sum(v);
  if (threadIdx.x == 0) {
    if (take_sqrt) v = sqrt(v);
    out[(size_t)r * stride + off] = v;
    if (copy_to) copy_to[r] = v;
  }
}

// This lane's share (entries lane, lane + 64, ...) of the n partials p, four
// loads in flight; the grouping depends on n only, so the sum is deterministic.
__device__ __forceinline__ double lane_sum4(const double* __restrict__ p, int n, int lane) {
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
  int q = lane;
  for (; q + 192 < n; q += 256) {
    a0 += p[q];
    a1 += p[q + 64];
    a2 += p[q + 128];
    a3 += p[q + 192];
  }
  for (; q < n; q += 64) a0 += p[q];
  return (a0 + a1) + (a2 + a3);
}

// w <- w - a_j v_j - b_{j-1} v_{j-1}; partial |w|^2 out.  a_j = v_j.w is
// reduced by every wavefront from its replica's per-wave partials in wave order
// (a wavefront's 64 slots always belong to one replica), so all wavefronts of a
// replica use the identical value.
__global__ void k_lz_orth(const RepDev* __restrict__ reps, const int* __restrict__ slot_rep, int slot_begin,
                          int nslots, double* w, const double* vj, const double* vjm1, double* scal, int j,
                          const int* done, const double* partial_alpha, const double* cluster_alpha,
                          double* partial) {
  const int lane = threadIdx.x & 63;
  int s = slot_begin + blockIdx.x * blockDim.x + threadIdx.x;
  int s_w = s - lane;  // first slot of this wavefront
  if (s_w >= slot_begin + nslots) return;
  int r = slot_rep[s_w];
  if (r < 0) return;
  const RepDev& R = reps[r];
  if (!R.active || (done && done[r])) return;
  // alpha from the partials the cluster pass wrote (one per bead cluster) or the
  // per-wavefront ones of k_wave_dot
  double a = cluster_alpha ? lane_sum4(cluster_alpha + R.cl0, R.ncl, lane)
                           : lane_sum4(partial_alpha + R.wave0, R.nwaves, lane);
  a = warp_sum(a);
  double* sc = scal + (size_t)r * SCAL_STRIDE;
  if (lane == 0 && s_w == R.slot0) sc[S_AL + j - 1] = a;
  double v = 0.0;
  if (s - R.slot0 < R.n) {
    double b = (j > 1) ? sc[S_BE + j - 2] : 0.0;
    for (int c = 0; c < 3; ++c) {
      double x = w[3 * s + c] - a * vj[3 * s + c];
      if (j > 1) x -= b * vjm1[3 * s + c];
      w[3 * s + c] = x;
      v += x * x;
    }
  }
  v = warp_sum(v);
  if (lane == 0) partial[s >> 6] = v;
}

// b_j from the partials, then c_j = b0 T_j^(1/2) e1 and the stopping test
// |c_j - [c_{j-1}; 0]| < tol |c_j|.  One workgroup of LZ_TRI_TH threads per replica.
//
// T^(1/2) e1 = T T^(-1/2) e1 with T^(-1/2) = (2/pi) int_R e^s (T + e^(2s) I)^(-1) ds,
// the substitution t = e^s of the classical integral (2/pi) int_0^inf (T + t^2)^(-1) dt.
// For one eigenvalue the integrand is lambda^(-1/2) sech(s - ln(lambda)/2)/pi, so
// the trapezoid rule with step h has relative error ~4 exp(-pi^2/h) for every
// eigenvalue alike (poles of sech at +-i pi/2), and truncating at U beyond the
// spectrum costs ~(2/pi) e^-U.  h = 0.5, U = 20 give ~1e-8, two orders below the
// stopping tolerance.  The spectrum is bracketed by Gershgorin above; the
// lower end of the node range is 1e-12 of that bound (condition numbers of the
// mobility are far below 1e12), which takes 109 nodes.  Each node is one SPD
// tridiagonal solve (T + sigma I) x = e1 by one thread: backward pivots, storing
// the ratios x_(i+1)/x_i = -e_i / pivot_(i+1), then x forwards by those ratios.
// The reciprocals are v_rcp_f64 refined by one Newton step (relative error
// ~1e-15 per pivot).  Columns are laid out [i][thread] with a row stride of
// LZ_TRI_TH + 1 doubles, so that the row sums (one thread per row, four
// interleaved partial sums over the nodes in fixed order) do not all fall on
// one LDS bank; the result does not depend on scheduling.
#define LZ_QH 0.5
#define LZ_QU 20.0
#define LZ_TRI_TH 128

#define LZ_COLS (LZ_TRI_TH + 1)

__device__ __forceinline__ double rcp_nr1(double x) {
  const double y = __builtin_amdgcn_rcp(x);
  return fma(fma(-x, y, 1.0), y, y);
}

__global__ __launch_bounds__(LZ_TRI_TH) void k_lz_tri(const RepDev* __restrict__ reps, int rep_begin,
                                                      const double* partial, double* scal, int j, double tol,
                                                      int maxit, int* done, double* bdiv, int* iters) {
  int r = rep_begin + blockIdx.x;
  const RepDev& R = reps[r];
  if (!R.active || done[r]) return;
  const int t = threadIdx.x, lane = t & 63;
  const bool w0 = t < 64;
  double* sc = scal + (size_t)r * SCAL_STRIDE;
  __shared__ double d[LZ_MAXIT], e[LZ_MAXIT], e2[LZ_MAXIT], u[LZ_MAXIT];
  __shared__ double col[LZ_MAXIT * LZ_COLS];  // ratios, then weighted first columns: [i][thread]
  __shared__ double bnd;
  __shared__ int brk;
  if (w0) {
    double v = 0.0;
    for (int w = lane; w < R.nwaves; w += 64) v += partial[R.wave0 + w];
    v = warp_sum(v);
    const double bj = sqrt(v);
    if (lane == 0) {
      sc[S_BE + j - 1] = bj;
      bdiv[r] = bj;
      iters[r] = j;
      sc[S_IT] = (double)j;
      brk = !(bj > 1e-300);
    }
  }
  if (!(sc[S_B0] > 0.0)) {  // z = 0: y = 0 exactly
    if (t == 0) {
      sc[S_C] = 0.0;
      sc[S_ERR] = 0.0;
      done[r] = 1;
    }
    return;
  }
  __syncthreads();
  if (j < (int)sc[S_JMIN] && j < maxit && !brk) return;  // too early to stop: beta_j only
  const int n = j;
  for (int i = t; i < n; i += LZ_TRI_TH) {
    d[i] = sc[S_AL + i];
    e[i] = (i < n - 1) ? sc[S_BE + i] : 0.0;
    e2[i] = e[i] * e[i];
  }
  __syncthreads();
  if (w0) {  // Gershgorin upper bound of the spectrum
    double hi = 0.0;
    for (int i = lane; i < n; i += 64) hi = fmax(hi, d[i] + fabs(e[i]) + (i > 0 ? fabs(e[i - 1]) : 0.0));
    hi = warp_max(hi);
    if (lane == 0) bnd = hi;
  }
  __syncthreads();
  const double hb = fmax(bnd, 1e-290);
  const double slo = 0.5 * log(hb * 1e-12) - LZ_QU, shi = 0.5 * log(hb) + LZ_QU;
  const int nodes = min(LZ_TRI_TH, (int)ceil((shi - slo) / LZ_QH) + 1);  // 109
  if (t < nodes) {
    const double s = slo + t * LZ_QH;
    const double sig = exp(2.0 * s);
    const double wk = LZ_QH * 0.6366197723675814 * exp(s);
    double ip = rcp_nr1(d[n - 1] + sig);
    for (int i = n - 2; i >= 0; --i) {
      col[(i + 1) * LZ_COLS + t] = -e[i] * ip;
      ip = rcp_nr1(fma(-e2[i], ip, d[i] + sig));
    }
    double xi = ip;  // (T + sig)^-1 e1, first entry
    col[t] = wk * xi;
    int i = 1;
    for (; i + 3 < n; i += 4) {  // ratios loaded ahead of the product chain
      const double r0 = col[i * LZ_COLS + t], r1 = col[(i + 1) * LZ_COLS + t];
      const double r2 = col[(i + 2) * LZ_COLS + t], r3 = col[(i + 3) * LZ_COLS + t];
      xi *= r0;
      col[i * LZ_COLS + t] = wk * xi;
      xi *= r1;
      col[(i + 1) * LZ_COLS + t] = wk * xi;
      xi *= r2;
      col[(i + 2) * LZ_COLS + t] = wk * xi;
      xi *= r3;
      col[(i + 3) * LZ_COLS + t] = wk * xi;
    }
    for (; i < n; ++i) {
      xi *= col[i * LZ_COLS + t];
      col[i * LZ_COLS + t] = wk * xi;
    }
  }
  __syncthreads();
  for (int i = t; i < n; i += LZ_TRI_TH) {
    const double* row = col + i * LZ_COLS;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    int q = 0;
    for (; q + 3 < nodes; q += 4) {
      a0 += row[q];
      a1 += row[q + 1];
      a2 += row[q + 2];
      a3 += row[q + 3];
    }
    for (; q < nodes; ++q) a0 += row[q];
    u[i] = (a0 + a1) + (a2 + a3);
  }
  __syncthreads();
  if (!w0) return;
  const double b0 = sc[S_B0];
  const int nprev = (int)sc[S_ERR + 1];
  double num = 0.0, den = 0.0;
  for (int i = lane; i < n; i += 64) {
    double ci = d[i] * u[i];
    if (i > 0) ci += e[i - 1] * u[i - 1];
    if (i < n - 1) ci += e[i] * u[i + 1];
    ci *= b0;
    double cp = (i < nprev) ? sc[S_C + i] : 0.0;
    num += (ci - cp) * (ci - cp);
    den += ci * ci;
    sc[S_CP + i] = ci;
  }
  num = warp_sum(num);
  den = warp_sum(den);
  for (int i = lane; i < n; i += 64) sc[S_C + i] = sc[S_CP + i];
  double err = (nprev > 0) ? sqrt(num / den) : 1.0;
  if (lane == 0) {
    sc[S_ERR] = err;
    sc[S_ERR + 1] = (double)n;
    bool conv = nprev > 0 && err < tol;
    if (conv || j >= maxit || brk) done[r] = conv ? 1 : 2;
  }
}

// y = V c (first iters[r] basis vectors), times yscale
__global__ void k_lz_combine(const RepDev* __restrict__ reps, const int* __restrict__ slot_rep, int slot_begin,
                             int nslots, const double* V, int nst, const double* scal, const int* iters, double* y) {
  int s = slot_begin + blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= slot_begin + nslots) return;
  int r = slot_rep[s];
  if (r < 0) return;
  const RepDev& R = reps[r];
  if (!R.active) return;
  if (s - R.slot0 >= R.n) {
    y[3 * s] = y[3 * s + 1] = y[3 * s + 2] = 0.0;
    return;
  }
  const double* sc = scal + (size_t)r * SCAL_STRIDE;
  int it = iters[r];
  double a0 = 0, a1 = 0, a2 = 0;
  for (int k = 0; k < it; ++k) {
    const double* vk = V + (size_t)k * 3 * nst;
    double ck = sc[S_C + k];
    a0 += ck * vk[3 * s];
    a1 += ck * vk[3 * s + 1];
    a2 += ck * vk[3 * s + 2];
  }
  y[3 * s] = a0;
  y[3 * s + 1] = a1;
  y[3 * s + 2] = a2;
}

// TEA: y_i = C_i (.) (beta (M z)_i + (1 - beta) diag_i z_i)
__global__ void k_tea(const int* __restrict__ slot_rep, int slot_begin, int nslots, const double* mz, const double* z,
                      const double* diag, const double* beta, const double* C, double* y) {
  int s = slot_begin + blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= slot_begin + nslots) return;
  int r = slot_rep[s];
  if (r < 0) return;
  double b = beta[r];
  for (int c = 0; c < 3; ++c)
    y[3 * s + c] = C[3 * s + c] * (b * mz[3 * s + c] + (1.0 - b) * diag[3 * r + c] * z[3 * s + c]);
}

// ------------------------------------------------------------------ launchers
void lz_wave_dot(hipStream_t st, const RepDev* reps, const int* slot_rep, int slot_begin, int nslots, const double* x,
                 const double* y, const int* done, double* partial) {
  if (nslots > 0)
    k_wave_dot<<<(nslots + 255) / 256, 256, 0, st>>>(reps, slot_rep, slot_begin, nslots, x, y, done, partial);
}
void lz_rep_reduce(hipStream_t st, const RepDev* reps, int rep_begin, int nrep, const double* partial,
                   const int* done, double* out, int stride, int off, int take_sqrt, double* copy_to) {
  if (nrep > 0)
    k_rep_reduce<<<nrep, 64, 0, st>>>(reps, rep_begin, partial, done, out, stride, off, take_sqrt, copy_to);
}
void lz_orth(hipStream_t st, const RepDev* reps, const int* slot_rep, int slot_begin, int nslots, double* w,
             const double* vj, const double* vjm1, double* scal, int j, const int* done, const double* partial_alpha,
             const double* cluster_alpha, double* partial) {
  if (nslots > 0)
    k_lz_orth<<<(nslots + 255) / 256, 256, 0, st>>>(reps, slot_rep, slot_begin, nslots, w, vj, vjm1, scal, j, done,
                                                     partial_alpha, cluster_alpha, partial);
}
void lz_tri(hipStream_t st, const RepDev* reps, int rep_begin, int nrep, const double* partial, double* scal, int j,
            double tol, int maxit, int* done, double* bdiv, int* iters) {
  if (nrep > 0) k_lz_tri<<<nrep, LZ_TRI_TH, 0, st>>>(reps, rep_begin, partial, scal, j, tol, maxit, done, bdiv, iters);
}
void lz_combine(hipStream_t st, const RepDev* reps, const int* slot_rep, int slot_begin, int nslots, const double* V,
                int nst, const double* scal, const int* iters, double* y) {
  if (nslots > 0)
    k_lz_combine<<<(nslots + 255) / 256, 256, 0, st>>>(reps, slot_rep, slot_begin, nslots, V, nst, scal, iters, y);
}
void lz_tea(hipStream_t st, const int* slot_rep, int slot_begin, int nslots, const double* mz, const double* z,
            const double* diag, const double* beta, const double* C, double* y) {
  if (nslots > 0)
    k_tea<<<(nslots + 255) / 256, 256, 0, st>>>(slot_rep, slot_begin, nslots, mz, z, diag, beta, C, y);
}

// zero the "previous check" length before a new solve
// Start of a solve: no previous c, and the square root is first evaluated at half
// the iteration count the replica's previous solve stopped at (at least 2): a
// solve converges within a few percent of its predecessor, so the earlier
// evaluations could not meet the stopping test.  One that converges sooner stops
// at the first evaluated pair, a few iterations late.
__global__ void k_lz_reset(double* scal, int rep_begin, int nrep) {
  int r = rep_begin + blockIdx.x * blockDim.x + threadIdx.x;
  if (r < rep_begin + nrep) {
    double* sc = scal + (size_t)r * SCAL_STRIDE;
    sc[S_ERR + 1] = 0.0;
    sc[S_JMIN] = fmax(2.0, floor(0.5 * sc[S_IT]));
  }
}
void lz_reset_checks(hipStream_t st, double* scal, int rep_begin, int nrep) {
  if (nrep > 0) k_lz_reset<<<(nrep + 63) / 64, 64, 0, st>>>(scal, rep_begin, nrep);
}
