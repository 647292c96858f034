// Conservative forces, FP64, gathered per bead (no atomics, deterministic):
// springs + bending (bondforce, S/main26.c:541-618), Lennard-Jones over the
// Verlet list (LJforce, S/main26.c:619-681; extensional form X/main14.c:659)
// and all-pairs screened Coulomb between charged beads (coulombForce,
// S/main26.c:1891-1925).  Pair vectors come from the reference's minimum image
// and stay in the wrap frame (rotation to the lab frame is done on copy-out).
#include "common.h"

__device__ __forceinline__ double spring_fs(const RepDev& R, double rr, int* err) {
  double qmax2 = R.qmax * R.qmax;
  if (R.spring_fene) {
    if (rr > qmax2) {
      atomicOr(err, 1);
      return 0.0;
    }
    return -R.kappas / (1.0 - rr / qmax2);
  }
  double r = sqrt(rr);
  return -R.kappas * (r - R.qmax) / r;
}

__device__ __forceinline__ void nid_slots(const RepDev& R, const double* rx, const double* ry, const double* rz,
                                          int a, int b, double d[3]) {
  ref_nid(R, rx[a], ry[a], rz[a], rx[b], ry[b], rz[b], d[0], d[1], d[2], false);
}

// bending contribution to bead i from the angle centred at bead c (same chain,
// 1 <= position(c) <= nb-2)
__device__ __forceinline__ void bend_term(const RepDev& R, const double* rx, const double* ry, const double* rz, int c,
                                          int i, double F[3]) {
  double n1[3], n2[3];
  nid_slots(R, rx, ry, rz, c, c - 1, n1);
  nid_slots(R, rx, ry, rz, c, c + 1, n2);
  double a1 = sqrt(n1[0] * n1[0] + n1[1] * n1[1] + n1[2] * n1[2]);
  double a2 = sqrt(n2[0] * n2[0] + n2[1] * n2[1] + n2[2] * n2[2]);
  double var = (n1[0] * n2[0] + n1[1] * n2[1] + n1[2] * n2[2]) / (a1 * a2);
  double thetab = M_PI - acos(var);
  double coeff = -R.kappab * thetab / sqrt(1 - var * var);
  for (int k = 0; k < 3; ++k) {
    if (i == c)
      F[k] += coeff * ((n1[k] + n2[k]) / (a1 * a2) - var * (n1[k] / (a1 * a1) + n2[k] / (a2 * a2)));
    else if (i == c - 1)
      F[k] += coeff * (var * n1[k] / (a1 * a1) - n2[k] / (a1 * a2));
    else
      F[k] += coeff * (var * n2[k] / (a2 * a2) - n1[k] / (a1 * a2));
  }
}

__global__ void k_forces(const RepDev* __restrict__ reps, const int* __restrict__ slot_rep, int slot_begin,
                         int nslots, const double* __restrict__ rx, const double* __restrict__ ry,
                         const double* __restrict__ rz, int nst, const int* __restrict__ nlj,
                         const int* __restrict__ llj, double* f, int* err) {
  int i = slot_begin + blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= slot_begin + nslots) return;
  int r = slot_rep[i];
  if (r < 0) return;
  const RepDev& R = reps[r];
  if (!R.active) return;
  int li = i - R.slot0;
  if (li >= R.n) {
    f[3 * i] = f[3 * i + 1] = f[3 * i + 2] = 0.0;
    return;
  }
  int pos = li % R.nb;
  double F[3] = {0.0, 0.0, 0.0}, d[3];
  if (pos > 0) {  // spring (i, i-1): f_i += Fs * (r_i - r_{i-1})
    nid_slots(R, rx, ry, rz, i, i - 1, d);
    double fs = spring_fs(R, d[0] * d[0] + d[1] * d[1] + d[2] * d[2], err);
    for (int k = 0; k < 3; ++k) F[k] += fs * d[k];
  }
  if (pos < R.nb - 1) {  // spring (i+1, i): f_i -= Fs * (r_{i+1} - r_i)
    nid_slots(R, rx, ry, rz, i + 1, i, d);
    double fs = spring_fs(R, d[0] * d[0] + d[1] * d[1] + d[2] * d[2], err);
    for (int k = 0; k < 3; ++k) F[k] -= fs * d[k];
  }
  if (R.kappab > 0.0) {
    for (int c = i - 1; c <= i + 1; ++c) {
      int pc = pos + (c - i);
      if (pc >= 1 && pc <= R.nb - 2) bend_term(R, rx, ry, rz, c, i, F);
    }
  }
  int m = nlj[i];
  double xi = rx[i], yi = ry[i], zi = rz[i];
  for (int k = 0; k < m; ++k) {
    int j = llj[(size_t)k * nst + i];
    ref_nid(R, xi, yi, zi, rx[j], ry[j], rz[j], d[0], d[1], d[2], false);
    double rr = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    if (rr < R.r2cut) {
      double ratio = 4.00 / rr;
      double r6 = ratio * ratio * ratio;
      double coeff = R.lj_ext ? (48.0 * R.epsilon / rr) * (r6 * r6 - 0.5 * r6) : (12 * R.epsilon / rr) * (r6 * r6 - r6);
      for (int q = 0; q < 3; ++q) F[q] += coeff * d[q];
    }
  }
  f[3 * i] = F[0];
  f[3 * i + 1] = F[1];
  f[3 * i + 2] = F[2];
}

void launch_forces(hipStream_t st, const RepDev* reps, const int* slot_rep, int slot_begin, int nslots,
                   const double* rx, const double* ry, const double* rz, int nslots_total, const int* nlj,
                   const int* llj, double* f, int* err) {
  if (nslots <= 0) return;
  k_forces<<<(nslots + 127) / 128, 128, 0, st>>>(reps, slot_rep, slot_begin, nslots, rx, ry, rz, nslots_total, nlj,
                                                 llj, f, err);
}

// Screened Coulomb between all charged beads of a replica: one wavefront per
// charge, lanes stride over the partners, FP64 sums combined across the
// wavefront.  qtrack holds bead indices local to the replica (Charge_track).
__global__ __launch_bounds__(256) void k_coulomb(const RepDev* __restrict__ reps, int rep_begin,
                                                 const double* __restrict__ rx, const double* __restrict__ ry,
                                                 const double* __restrict__ rz, const int* __restrict__ qtrack,
                                                 const int* __restrict__ qsign, double* f) {
  const RepDev& R = reps[rep_begin + blockIdx.y];
  if (!R.active) return;
  const int lane = threadIdx.x & 63;
  const int a = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int nq = R.nq;
  if (a >= nq) return;
  const int ia = R.slot0 + qtrack[R.q0 + a];
  const double xa = rx[ia], ya = ry[ia], za = rz[ia];
  const int qa = qsign[R.q0 + a];
  double F[3] = {0, 0, 0};
  for (int b = lane; b < nq; b += 64) {
    if (b == a) continue;
    int ib = R.slot0 + qtrack[R.q0 + b];
    double d[3];
    ref_nid(R, xa, ya, za, rx[ib], ry[ib], rz[ib], d[0], d[1], d[2], false);
    double rr = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    double rd = sqrt(rr);
    double Fc = R.lambda_b * (double)(qa * qsign[R.q0 + b]) * exp(-R.kappa_d * rd) * (1 + rd * R.kappa_d) / rr;
    for (int q = 0; q < 3; ++q) F[q] += Fc * d[q] / rd;
  }
  double fx = warp_sum(F[0]), fy = warp_sum(F[1]), fz = warp_sum(F[2]);
  if (lane == 0) {
    f[3 * ia] += fx;
    f[3 * ia + 1] += fy;
    f[3 * ia + 2] += fz;
  }
}

void launch_coulomb(hipStream_t st, const RepDev* reps, int rep_begin, int nrep, const double* rx, const double* ry,
                    const double* rz, const int* qtrack, const int* qsign, double* f, int max_nq) {
  if (nrep <= 0 || max_nq <= 0) return;
  dim3 grid((max_nq + 3) / 4, nrep);
  k_coulomb<<<grid, 256, 0, st>>>(reps, rep_begin, rx, ry, rz, qtrack, qsign, f);
}
