// One sample of the conformation-averaged TEA parameters: the pair sums of
// ewaldBin (S/main26.c:975-1260, X/main14.c:1009-1383) on the device.  Every
// bead i of a replica gets one workgroup; its threads take the partners j != i
// in turn and evaluate the reference's Ewald pair block Dt(i, j) (real-space
// images within rce with the reference's r <= 2 branch, plus the k-sum with
// cos(k.r) of the raw difference).  Per bead the workgroup returns
//   eps_i    = sum_{j != i} sum_ab Dt_ab / Dtr_a
//   rowsum_ia = sum_{j != i} sum_b (Dt_ab / Dtr_a)^2
// summed per thread in partner order and across threads in a fixed order, so
// the result does not depend on scheduling.  Each unordered pair is evaluated
// from both ends (Dt(j, i) = Dt(i, j)), which doubles the arithmetic but needs
// no scatter.  The k table (rk, M2) and Dtr come from the host.
#include "common.h"

// Real-space coefficients of the reference's pair block (S/main26.c:1150-1170):
// the Beenakker far form for r > 2, the reference's r <= 2 branch otherwise.
__device__ __forceinline__ void ref_real_coeffs(double r, double rr, const TeaRep& R, double& t1, double& t2) {
  const double a = R.alpha, a3 = R.alpha3, a5 = R.alpha5, a7 = R.alpha7;
  double C1, C2, C3, C4;
  if (r > 2.0) {
    C1 = 0.75 / r + 0.5 / (r * rr);
    C2 = 4.0 * a7 * rr * rr + 3.0 * a3 * rr - 20.0 * a5 * rr - 4.5 * a + 14.0 * a3 + a / rr;
    C3 = 0.75 / r - 1.5 / (r * rr);
    C4 = -4.0 * a7 * rr * rr - 3.0 * a3 * rr + 16.0 * a5 * rr + 1.5 * a - 2.0 * a3 - 3.0 * a / rr;
  } else {
    C1 = 1.0 - 0.28125 * r;
    C2 = 64.0 * a7 + 12.0 * a3 - 80.0 * a5 - 4.5 * a + 14.0 * a3 + 0.25 * a;
    C3 = 0.09375 * r;
    C4 = -64.0 * a7 - 12.0 * a3 + 64.0 * a5 + 1.5 * a - 2.0 * a3 - 0.75 * a;
  }
  const double ec = erfc(a * r), g = exp(-R.alpha2 * rr) * 0.56418958354775628695;  // 1/sqrt(pi)
  t1 = C1 * ec + C2 * g;
  t2 = C3 * ec + C4 * g;
}

// The reference's Ewald pair block Dt for the raw difference (dxo, dyo, dzo):
// every image within rce with ref_real_coeffs, then the k-sum with cos(k.r) of
// the raw difference (S/main26.c:1150-1219, S/gridHI.c:148-244).  kt holds 12
// doubles per wave vector (rkx, rky, rkz, M2[9]).
__device__ __forceinline__ void tea_pair_block(const TeaRep& R, const double* kt, double dxo, double dyo, double dzo,
                                               double Dt[9]) {
#pragma unroll
  for (int q = 0; q < 9; ++q) Dt[q] = 0.0;
  for (int nx = -R.nmax; nx <= R.nmax; ++nx)
    for (int ny = -R.nmax; ny <= R.nmax; ++ny)
      for (int nz = -R.nmax; nz <= R.nmax; ++nz) {
        const double dx = dxo + nx * R.L1[0] + ny * R.L2[0], dy = dyo + nx * R.L1[1] + ny * R.L2[1],
                     dz = dzo + nz * R.L;
        const double rr = dx * dx + dy * dy + dz * dz, rd = sqrt(rr);
        if (rd < R.rce) {
          double t1, t2;
          ref_real_coeffs(rd, rr, R, t1, t2);
          const double d[3] = {dx, dy, dz};
#pragma unroll
          for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b) Dt[3 * a + b] += (a == b ? t1 : 0.0) + t2 * d[a] * d[b] / rr;
        }
      }
  for (int q = 0; q < R.nk; ++q) {
    const double* k = kt + 12 * q;
    const double c = cos(k[0] * dxo + k[1] * dyo + k[2] * dzo);
#pragma unroll
    for (int a = 0; a < 9; ++a) Dt[a] += k[3 + a] * c;
  }
}

__global__ __launch_bounds__(TEA_THREADS) void k_tea_rows(const TeaRep* __restrict__ reps, const int* __restrict__ bead_rep,
                                                          const double* __restrict__ rx, const double* __restrict__ ry,
                                                          const double* __restrict__ rz, const double* __restrict__ ktab,
                                                          double* __restrict__ eps_out, double* __restrict__ rowsum_out) {
  __shared__ double kt[TEA_MAXK * 12];
  __shared__ double red[TEA_THREADS / 64][4];
  const int r = bead_rep[blockIdx.x];
  const TeaRep& R = reps[r];
  const int i = R.slot0 + (blockIdx.x - R.bead0);
  const int t = threadIdx.x;
  for (int q = t; q < 12 * R.nk; q += TEA_THREADS) kt[q] = ktab[12 * (size_t)R.k0 + q];
  __syncthreads();
  const double xi = rx[i], yi = ry[i], zi = rz[i];
  double e = 0.0, s0 = 0.0, s1 = 0.0, s2 = 0.0;
  for (int jj = t; jj < R.n; jj += TEA_THREADS) {
    const int j = R.slot0 + jj;
    if (j == i) continue;
    double Dt[9];
    tea_pair_block(R, kt, rx[j] - xi, ry[j] - yi, rz[j] - zi, Dt);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      double rs = 0.0;
#pragma unroll
      for (int b = 0; b < 3; ++b) {
        const double incr = Dt[3 * a + b] / R.dtr[a];
        e += incr;
        rs += incr * incr;
      }
      if (a == 0) s0 += rs;
      else if (a == 1) s1 += rs;
      else s2 += rs;
    }
  }
  e = warp_sum(e);
  s0 = warp_sum(s0);
  s1 = warp_sum(s1);
  s2 = warp_sum(s2);
  const int w = t >> 6;
  if ((t & 63) == 0) {
    red[w][0] = e;
    red[w][1] = s0;
    red[w][2] = s1;
    red[w][3] = s2;
  }
  __syncthreads();
  if (t < 4) {
    double v = 0.0;
    for (int q = 0; q < TEA_THREADS / 64; ++q) v += red[q][t];
    if (t == 0) eps_out[i] = v;
    else rowsum_out[3 * i + t - 1] = v;
  }
}

void launch_tea_rows(hipStream_t st, const TeaRep* reps, const int* bead_rep, int nbeads, const double* rx,
                     const double* ry, const double* rz, const double* ktab, double* eps_out, double* rowsum_out) {
  if (nbeads <= 0) return;
  k_tea_rows<<<nbeads, TEA_THREADS, 0, st>>>(reps, bead_rep, rx, ry, rz, ktab, eps_out, rowsum_out);
}

// Far-field table of the reference's gridHI (S/gridHI.c:3-244, X/gridHI.c): the
// pair block at every bin displacement of the table geometry, one thread per
// bin, one table per strain phase (blockIdx.y); the origin bin takes the
// reference's placeholder displacement (0.6, 0.6, 0.6).  The k table is read
// from global memory (it can exceed LDS at M_HI = 6); every thread of a
// wavefront reads the same entry.
__global__ __launch_bounds__(256) void k_far_table(const TeaRep* __restrict__ phases, const double* __restrict__ ktab,
                                                   int nbx, int nby, int nbz, double bsx, double bsy, double bsz,
                                                   double* __restrict__ table) {
  const TeaRep& R = phases[blockIdx.y];
  const int nbins = nbx * nby * nbz;
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nbins) return;
  const int nhx = (nbx - 1) / 2, nhy = (nby - 1) / 2, nhz = (nbz - 1) / 2;
  const int origin = nhx * nby * nbz + nhy * nbz + nhz;
  double dr[3] = {0.6, 0.6, 0.6};
  if (b != origin) {
    const int bx = b / (nby * nbz), rem = b - bx * nby * nbz, by = rem / nbz, bz = rem - by * nbz;
    dr[0] = (bx - nhx) * bsx;
    dr[1] = (by - nhy) * bsy;
    dr[2] = (bz - nhz) * bsz;
  }
  double Dt[9];
  tea_pair_block(R, ktab + 12 * (size_t)R.k0, dr[0], dr[1], dr[2], Dt);
  double* out = table + (size_t)blockIdx.y * 9 * nbins + 9 * (size_t)b;
#pragma unroll
  for (int a = 0; a < 9; ++a) out[a] = Dt[a];
}

void launch_far_table(hipStream_t st, const TeaRep* phases, int nphase, const double* ktab, const int nb[3],
                      const double bs[3], double* table) {
  const int nbins = nb[0] * nb[1] * nb[2];
  if (nbins <= 0 || nphase <= 0) return;
  dim3 grid((nbins + 255) / 256, nphase);
  k_far_table<<<grid, 256, 0, st>>>(phases, ktab, nb[0], nb[1], nb[2], bs[0], bs[1], bs[2], table);
}
