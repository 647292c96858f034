// Lanczos (Krylov) Brownian displacement y = M^(1/2) z per replica, FP64:
//   v_1 = z/|z|;  w = M v_j;  a_j = v_j.w;  w -= a_j v_j + b_{j-1} v_{j-1};  b_j = |w|
//   c_j = |z| T_j^(1/2) e_1 (T_j tridiagonal), stop when |c_j - [c_{j-1};0]| < tol |c_j|,
//   y = V_j c_j.
// This replaces the TEA product of updateHI (S/main26.c:787-804).  Reductions
// are two-level and ordered (per-wavefront partials, then one wavefront per
// replica), so results do not depend on scheduling.  The small tridiagonal
// eigenproblem is solved by implicit-shift QL on one wavefront per replica,
// each lane owning one row of the eigenvector matrix in LDS.
#include "common.h"

#define SCAL_STRIDE (4 * LZ_MAXIT + 8)
#define S_AL 0
#define S_BE LZ_MAXIT
#define S_C (2 * LZ_MAXIT)
#define S_CP (3 * LZ_MAXIT)
#define S_B0 (4 * LZ_MAXIT)
#define S_IT (4 * LZ_MAXIT + 1)
#define S_ERR (4 * LZ_MAXIT + 2)

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

// w <- w - a_j v_j - b_{j-1} v_{j-1}; partial |w|^2 out.  a_j = v_j.w is
// reduced by every wavefront from its replica's per-wave partials in wave order
// (a wavefront's 64 slots always belong to one replica), so all wavefronts of a
// replica use the identical value.
__global__ void k_lz_orth(const RepDev* __restrict__ reps, const int* __restrict__ slot_rep, int slot_begin,
                          int nslots, double* w, const double* vj, const double* vjm1, double* scal, int j,
                          const int* done, const double* partial_alpha, double* partial) {
  const int lane = threadIdx.x & 63;
  int s = slot_begin + blockIdx.x * blockDim.x + threadIdx.x;
  int s_w = s - lane;  // first slot of this wavefront
  if (s_w >= slot_begin + nslots) return;
  int r = slot_rep[s_w];
  if (r < 0) return;
  const RepDev& R = reps[r];
  if (!R.active || (done && done[r])) return;
  double a = 0.0;
  for (int q = lane; q < R.nwaves; q += 64) a += partial_alpha[R.wave0 + q];
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

// b_j from the partials; then, on check iterations (j even and >= 4, the last
// iteration, or a breakdown), c_j = b0 T_j^(1/2) e1 by implicit-shift QL and the
// stopping test |c_j - c_prev| < tol |c_j| against the previous check.  Block =
// one wavefront per replica; every lane runs the scalar QL recurrence on its
// own copy (registers / LDS written with identical values) and rotates its own
// row of the eigenvector matrix, so the recurrence needs no barriers.
__global__ __launch_bounds__(64) void k_lz_tri(const RepDev* __restrict__ reps, int rep_begin, const double* partial,
                                               double* scal, int j, double tol, int maxit, int* done, double* bdiv,
                                               int* iters) {
  int r = rep_begin + blockIdx.x;
  const RepDev& R = reps[r];
  if (!R.active || done[r]) return;
  const int lane = threadIdx.x;
  double* sc = scal + (size_t)r * SCAL_STRIDE;
  double v = 0.0;
  for (int w = lane; w < R.nwaves; w += 64) v += partial[R.wave0 + w];
  v = warp_sum(v);
  const double bj = sqrt(v);
  const bool breakdown = !(bj > 1e-300);
  const bool check = (j >= 4 && (j % 2) == 0) || j >= maxit || breakdown;
  if (lane == 0) {
    sc[S_BE + j - 1] = bj;
    bdiv[r] = bj;
    iters[r] = j;
    sc[S_IT] = (double)j;
  }
  if (!check) return;
  __shared__ double d[LZ_MAXIT], e[LZ_MAXIT];
  __shared__ double Q[LZ_MAXIT * 64];
  const int n = j;
  __syncthreads();
  if (lane < n) {
    d[lane] = sc[S_AL + lane];
    e[lane] = (lane < n - 1) ? sc[S_BE + lane] : 0.0;
  }
  for (int k = 0; k < n; ++k) Q[k * 64 + lane] = (k == lane) ? 1.0 : 0.0;
  __syncthreads();
  for (int l = 0; l < n; ++l) {
    for (int iter = 0; iter < 64; ++iter) {
      int m;
      for (m = l; m < n - 1; ++m) {
        double dd = fabs(d[m]) + fabs(d[m + 1]);
        if (fabs(e[m]) <= 2.220446049250313e-16 * dd) break;
      }
      if (m == l) break;
      double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
      double rr = sqrt(g * g + 1.0);
      g = d[m] - d[l] + e[l] / (g + copysign(rr, g));
      double s = 1.0, c = 1.0, p = 0.0;
      bool early = false;
      for (int i = m - 1; i >= l; --i) {
        double f = s * e[i], b = c * e[i];
        rr = sqrt(f * f + g * g);
        e[i + 1] = rr;
        if (rr == 0.0) {
          d[i + 1] -= p;
          e[m] = 0.0;
          early = true;
          break;
        }
        s = f / rr;
        c = g / rr;
        g = d[i + 1] - p;
        rr = (d[i] - g) * s + 2.0 * c * b;
        p = s * rr;
        d[i + 1] = g + p;
        g = c * rr - b;
        if (lane < n) {
          double qa = Q[(i + 1) * 64 + lane], qb = Q[i * 64 + lane];
          Q[(i + 1) * 64 + lane] = s * qb + c * qa;
          Q[i * 64 + lane] = c * qb - s * qa;
        }
      }
      if (early) continue;
      d[l] -= p;
      e[l] = g;
      e[m] = 0.0;
    }
  }
  __syncthreads();
  const double b0 = sc[S_B0];
  double ci = 0.0;
  if (lane < n) {
    for (int k = 0; k < n; ++k) ci += Q[k * 64 + lane] * sqrt(fmax(d[k], 0.0)) * Q[k * 64 + 0];
    ci *= b0;
  }
  const int nprev = (int)sc[S_ERR + 1];
  double cprev = (lane < nprev) ? sc[S_CP + lane] : 0.0;
  double num = warp_sum((ci - cprev) * (ci - cprev));
  double den = warp_sum(ci * ci);
  double err = (nprev > 0) ? sqrt(num / den) : 1.0;
  if (lane < n) {
    sc[S_C + lane] = ci;
    sc[S_CP + lane] = ci;
  }
  if (lane == 0) {
    sc[S_ERR] = err;
    sc[S_ERR + 1] = (double)n;
    if ((nprev > 0 && err < tol) || j >= maxit || breakdown) done[r] = (nprev > 0 && err < tol) ? 1 : 2;
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
             double* partial) {
  if (nslots > 0)
    k_lz_orth<<<(nslots + 255) / 256, 256, 0, st>>>(reps, slot_rep, slot_begin, nslots, w, vj, vjm1, scal, j, done,
                                                     partial_alpha, partial);
}
void lz_tri(hipStream_t st, const RepDev* reps, int rep_begin, int nrep, const double* partial, double* scal, int j,
            double tol, int maxit, int* done, double* bdiv, int* iters) {
  if (nrep > 0) k_lz_tri<<<nrep, 64, 0, st>>>(reps, rep_begin, partial, scal, j, tol, maxit, done, bdiv, iters);
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
__global__ void k_lz_reset(double* scal, int rep_begin, int nrep) {
  int r = rep_begin + blockIdx.x * blockDim.x + threadIdx.x;
  if (r < rep_begin + nrep) scal[(size_t)r * SCAL_STRIDE + S_ERR + 1] = 0.0;
}
void lz_reset_checks(hipStream_t st, double* scal, int rep_begin, int nrep) {
  if (nrep > 0) k_lz_reset<<<(nrep + 63) / 64, 64, 0, st>>>(scal, rep_begin, nrep);
}
