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

// Virial stress (the bondforce/LJforce tau accumulators).  The reference adds
// -2 x_a Fs x_b once per spring and per LJ pair (S/main26.c:576-578,662-664)
// with getNID's vector, which for KRBC is in the lab frame; gathered per bead,
// every pair is met from both ends, so each end adds -x_a Fs x_b.  Components
// xx, yy, xy.
__device__ __forceinline__ void lab_vec(const RepDev& R, const double d[3], double o[3]) {
#pragma clang fp contract(off)
  if (R.flow == 2) {
    o[0] = d[0] * R.ct - d[1] * R.st;
    o[1] = d[0] * R.st + d[1] * R.ct;
  } else {
    o[0] = d[0];
    o[1] = d[1];
  }
  o[2] = d[2];
}
__device__ __forceinline__ void vir_pair(const RepDev& R, double v[3], double fs, const double d[3]) {
  double l[3];
  lab_vec(R, d, l);
  v[0] -= l[0] * fs * l[0];
  v[1] -= l[1] * fs * l[1];
  v[2] -= l[0] * fs * l[1];
}
// The bending part of the virial, once per angle (centre c): the reference's
// eight terms (S/main26.c:606-613, X/main14.c:612-619), xx and yy only, two of
// the yy terms reading the z components.
__device__ __forceinline__ void bend_virial(const RepDev& R, const double* rx, const double* ry, const double* rz,
                                            int c, double v[3]) {
  double n1[3], n2[3], p[3], q[3];
  nid_slots(R, rx, ry, rz, c, c - 1, n1);
  nid_slots(R, rx, ry, rz, c, c + 1, n2);
  double a1 = sqrt(n1[0] * n1[0] + n1[1] * n1[1] + n1[2] * n1[2]);
  double a2 = sqrt(n2[0] * n2[0] + n2[1] * n2[1] + n2[2] * n2[2]);
  double var = (n1[0] * n2[0] + n1[1] * n2[1] + n1[2] * n2[2]) / (a1 * a2);
  double thetab = M_PI - acos(var);
  double coeff = -R.kappab * thetab / sqrt(1 - var * var);
  lab_vec(R, n1, p);
  lab_vec(R, n2, q);
  v[0] += p[0] * coeff * (var * p[0] / (a1 * a1) - q[0] / (a1 * a2));
  v[0] -= p[0] * coeff * ((p[0]) / (a1 * a2) - var * (p[0] / (a1 * a1)));
  v[0] += q[0] * coeff * ((q[0]) / (a1 * a2) - var * (q[0] / (a2 * a2)));
  v[0] -= q[0] * coeff * (var * q[0] / (a2 * a2) - p[0] / (a1 * a2));
  v[1] += p[1] * coeff * (var * p[1] / (a1 * a1) - q[1] / (a1 * a2));
  v[1] -= p[1] * coeff * ((p[2]) / (a1 * a2) - var * (p[2] / (a1 * a1)));
  v[1] += q[1] * coeff * ((q[2]) / (a1 * a2) - var * (q[2] / (a2 * a2)));
  v[1] -= q[1] * coeff * (var * q[1] / (a2 * a2) - p[1] / (a1 * a2));
}

__global__ void k_forces(const RepDev* __restrict__ reps, const int* __restrict__ slot_rep, int slot_begin,
                         int nslots, const double* __restrict__ rx, const double* __restrict__ ry,
                         const double* __restrict__ rz, int nst, const int* __restrict__ nlj,
                         const int* __restrict__ llj, double* f, double* vir, int* err) {
  int i = slot_begin + blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= slot_begin + nslots) return;
  int r = slot_rep[i];
  if (r < 0) return;
  const RepDev& R = reps[r];
  if (!R.active) return;
  int li = i - R.slot0;
  if (li >= R.n) {
    f[3 * i] = f[3 * i + 1] = f[3 * i + 2] = 0.0;
    if (vir) vir[3 * i] = vir[3 * i + 1] = vir[3 * i + 2] = 0.0;
    return;
  }
  int pos = li % R.nb;
  double F[3] = {0.0, 0.0, 0.0}, d[3], V[3] = {0.0, 0.0, 0.0};
  if (pos > 0) {  // spring (i, i-1): f_i += Fs * (r_i - r_{i-1})
    nid_slots(R, rx, ry, rz, i, i - 1, d);
    double fs = spring_fs(R, d[0] * d[0] + d[1] * d[1] + d[2] * d[2], err);
    for (int k = 0; k < 3; ++k) F[k] += fs * d[k];
    if (vir) vir_pair(R, V, fs, d);
  }
  if (pos < R.nb - 1) {  // spring (i+1, i): f_i -= Fs * (r_{i+1} - r_i)
    nid_slots(R, rx, ry, rz, i + 1, i, d);
    double fs = spring_fs(R, d[0] * d[0] + d[1] * d[1] + d[2] * d[2], err);
    for (int k = 0; k < 3; ++k) F[k] -= fs * d[k];
    if (vir) vir_pair(R, V, fs, d);
  }
  if (R.kappab > 0.0) {
    for (int c = i - 1; c <= i + 1; ++c) {
      int pc = pos + (c - i);
      if (pc >= 1 && pc <= R.nb - 2) bend_term(R, rx, ry, rz, c, i, F);
    }
    if (vir && pos >= 1 && pos <= R.nb - 2) bend_virial(R, rx, ry, rz, i, V);
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
      if (vir) vir_pair(R, V, coeff, d);
    }
  }
  f[3 * i] = F[0];
  f[3 * i + 1] = F[1];
  f[3 * i + 2] = F[2];
  if (vir) {
    vir[3 * i] = V[0];
    vir[3 * i + 1] = V[1];
    vir[3 * i + 2] = V[2];
  }
}

// tau[r] += the replica's per-bead virial terms, summed in a fixed order (one
// workgroup per replica: strided per-thread sums, then a tree in LDS).
#define VIR_TH 1024
__global__ __launch_bounds__(VIR_TH) void k_virial_acc(const RepDev* __restrict__ reps, int rep_begin,
                                                       const double* __restrict__ vir, double* tau) {
  const int r = rep_begin + blockIdx.x, t = threadIdx.x;
  const RepDev& R = reps[r];
  if (!R.active) return;
  __shared__ double red[3][VIR_TH];
  double a0 = 0.0, a1 = 0.0, a2 = 0.0;
  for (int s = R.slot0 + t; s < R.slot0 + R.n; s += VIR_TH) {
    a0 += vir[3 * s];
    a1 += vir[3 * s + 1];
    a2 += vir[3 * s + 2];
  }
  red[0][t] = a0;
  red[1][t] = a1;
  red[2][t] = a2;
  __syncthreads();
  for (int w = VIR_TH / 2; w > 0; w >>= 1) {
    if (t < w)
      for (int k = 0; k < 3; ++k) red[k][t] += red[k][t + w];
    __syncthreads();
  }
  if (t == 0)
    for (int k = 0; k < 3; ++k) tau[3 * r + k] += red[k][0];
}

void launch_virial_acc(hipStream_t st, const RepDev* reps, int rep_begin, int nrep, const double* vir, double* tau) {
  if (nrep > 0) k_virial_acc<<<nrep, VIR_TH, 0, st>>>(reps, rep_begin, vir, tau);
}

// Per-chain observables (calcExt, calcCOM, calcRg; S/main26.c:74-93,925-962,
// 2391-2415), one thread per chain in the reference's sequential order:
// p_{j+1} = p_j + getNID(j, j+1) from p_0 = 0 (lab frame), ext = max - min of
// p_x, com = mean p, rg = mean (p - com)(p - com) as xx, xy, xz, yy, yz, zz.
// The unwrapped chain is walked twice (sums, then the tensor about the mean);
// no contraction, so the values equal a sequential FP64 evaluation.
__global__ void k_chain_obs(const RepDev* __restrict__ reps, int rep_begin, const int* __restrict__ chain0,
                            const double* __restrict__ rx, const double* __restrict__ ry,
                            const double* __restrict__ rz, double* ext, double* com, double* rg) {
#pragma clang fp contract(off)
  const int r = rep_begin + blockIdx.y;
  const RepDev& R = reps[r];
  const int ch = blockIdx.x * blockDim.x + threadIdx.x;
  if (!R.active || ch >= R.n / R.nb) return;
  const int b0 = R.slot0 + ch * R.nb, k = chain0[r] + ch;
  double px = 0.0, py = 0.0, pz = 0.0, mx = 0.0, mn = 0.0, cx = 0.0, cy = 0.0, cz = 0.0;
  for (int j = 0; j < R.nb; ++j) {
    if (j > 0) {
      double d[3];
      ref_nid(R, rx[b0 + j - 1], ry[b0 + j - 1], rz[b0 + j - 1], rx[b0 + j], ry[b0 + j], rz[b0 + j], d[0], d[1],
              d[2], true);
      px += d[0];
      py += d[1];
      pz += d[2];
    }
    mx = j ? fmax(mx, px) : px;
    mn = j ? fmin(mn, px) : px;
    cx += px;
    cy += py;
    cz += pz;
  }
  cx /= R.nb;
  cy /= R.nb;
  cz /= R.nb;
  double g[6] = {0, 0, 0, 0, 0, 0};
  px = py = pz = 0.0;
  for (int j = 0; j < R.nb; ++j) {
    if (j > 0) {
      double d[3];
      ref_nid(R, rx[b0 + j - 1], ry[b0 + j - 1], rz[b0 + j - 1], rx[b0 + j], ry[b0 + j], rz[b0 + j], d[0], d[1],
              d[2], true);
      px += d[0];
      py += d[1];
      pz += d[2];
    }
    const double a = px - cx, b = py - cy, e = pz - cz;
    g[0] += a * a;
    g[1] += a * b;
    g[2] += a * e;
    g[3] += b * b;
    g[4] += b * e;
    g[5] += e * e;
  }
  ext[k] = mx - mn;
  com[3 * k] = cx;
  com[3 * k + 1] = cy;
  com[3 * k + 2] = cz;
  for (int q = 0; q < 6; ++q) rg[6 * k + q] = g[q] / R.nb;
}

// The charge displacement log of updateHI/updateFD (S/main26.c:815-838,
// 2436-2446; X/main14.c:853-876).  At the start of a step each bead's offset
// from its chain's unwrapped centre of mass, rel_j = com - p_j (lab frame, the
// sequential unwrapping of calcCOM), is stored; after the position update every
// charge adds its bead's step displacement without the flow (dt M F + noise
// term, lab frame) plus the flow taken relative to the centre of mass: shear
// dx += dt edot rel_y; planar extension dx += dt edot rel_x, dy -= dt edot rel_y.
// (The unwrapping here is a prefix sum, so rel may differ from the sequential
// walk in the last bits.)
__global__ __launch_bounds__(256) void k_charge_rel(const RepDev* __restrict__ reps, int rep_begin,
                                                    const double* __restrict__ rx, const double* __restrict__ ry,
                                                    const double* __restrict__ rz, double* rel) {
  // one wavefront per chain: lane j takes getNID(j-1, j) of its 64-bead block and
  // a shuffle prefix sum (plus the running carry) gives p_j
  const RepDev& R = reps[rep_begin + blockIdx.y];
  const int lane = threadIdx.x & 63, ch = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (!R.active || R.nq == 0 || ch >= R.n / R.nb) return;
  const int b0 = R.slot0 + ch * R.nb;
  double carry[3] = {0.0, 0.0, 0.0}, sum[3] = {0.0, 0.0, 0.0};
  for (int base = 0; base < R.nb; base += 64) {
    const int j = base + lane;
    double d[3] = {0.0, 0.0, 0.0};
    if (j > 0 && j < R.nb)
      ref_nid(R, rx[b0 + j - 1], ry[b0 + j - 1], rz[b0 + j - 1], rx[b0 + j], ry[b0 + j], rz[b0 + j], d[0], d[1],
              d[2], true);
    for (int o = 1; o < 64; o <<= 1)
      for (int k = 0; k < 3; ++k) {
        const double u = __shfl_up(d[k], o, 64);
        if (lane >= o) d[k] += u;
      }
    for (int k = 0; k < 3; ++k) {
      const double pj = carry[k] + d[k];
      if (j < R.nb) {
        rel[3 * (b0 + j) + k] = pj;
        sum[k] += pj;
      }
      carry[k] += __shfl(d[k], 63, 64);
    }
  }
  double com[3];
  for (int k = 0; k < 3; ++k) com[k] = warp_sum(sum[k]) / R.nb;
  for (int j = lane; j < R.nb; j += 64)
    for (int k = 0; k < 3; ++k) rel[3 * (b0 + j) + k] = com[k] - rel[3 * (b0 + j) + k];
}

__global__ void k_charge_log(const RepDev* __restrict__ reps, int rep_begin, const int* __restrict__ qtrack,
                             const double* __restrict__ mf, const double* __restrict__ bz,
                             const double* __restrict__ rel, double* qdisp) {
#pragma clang fp contract(off)
  const RepDev& R = reps[rep_begin + blockIdx.y];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (!R.active || i >= R.nq) return;
  const int s = R.slot0 + qtrack[R.q0 + i];
  double d[3];
  for (int c = 0; c < 3; ++c) d[c] = R.dt * mf[3 * s + c] + R.sqrt2dt * bz[3 * s + c];
  double dx = d[0], dy = d[1];
  if (R.flow == 2) {  // the update's rotation to the lab frame
    dx = d[0] * R.ct - d[1] * R.st;
    dy = d[0] * R.st + d[1] * R.ct;
    dx += R.dt * R.edot * rel[3 * s];
    dy -= R.dt * R.edot * rel[3 * s + 1];
  } else {
    dx += R.dt * R.edot * rel[3 * s + 1];
  }
  double* o = qdisp + 3 * (R.q0 + i);
  o[0] += dx;
  o[1] += dy;
  o[2] += d[2];
}

void launch_charge_rel(hipStream_t st, const RepDev* reps, int rep_begin, int nrep, int max_chains, const double* rx,
                       const double* ry, const double* rz, double* rel) {
  if (nrep <= 0 || max_chains <= 0) return;
  k_charge_rel<<<dim3((max_chains + 3) / 4, nrep), 256, 0, st>>>(reps, rep_begin, rx, ry, rz, rel);
}

void launch_charge_log(hipStream_t st, const RepDev* reps, int rep_begin, int nrep, int max_nq, const int* qtrack,
                       const double* mf, const double* bz, const double* rel, double* qdisp) {
  if (nrep <= 0 || max_nq <= 0) return;
  k_charge_log<<<dim3((max_nq + 255) / 256, nrep), 256, 0, st>>>(reps, rep_begin, qtrack, mf, bz, rel, qdisp);
}

void launch_chain_obs(hipStream_t st, const RepDev* reps, int rep_begin, int nrep, int max_chains, const int* chain0,
                      const double* rx, const double* ry, const double* rz, double* ext, double* com, double* rg) {
  if (nrep <= 0 || max_chains <= 0) return;
  dim3 grid((max_chains + 63) / 64, nrep);
  k_chain_obs<<<grid, 64, 0, st>>>(reps, rep_begin, chain0, rx, ry, rz, ext, com, rg);
}

void launch_forces(hipStream_t st, const RepDev* reps, const int* slot_rep, int slot_begin, int nslots,
                   const double* rx, const double* ry, const double* rz, int nslots_total, const int* nlj,
                   const int* llj, double* f, double* vir, int* err) {
  if (nslots <= 0) return;
  k_forces<<<(nslots + 127) / 128, 128, 0, st>>>(reps, slot_rep, slot_begin, nslots, rx, ry, rz, nslots_total, nlj,
                                                 llj, f, vir, err);
}

// Screened Coulomb between all charged beads of a replica.  Each wavefront
// holds COUL_PER_WAVE charges and its lanes stride over the partners; the
// workgroup stages the partners' positions and signs in LDS tile by tile
// (gathered through qtrack once per workgroup instead of once per charge), so
// the pair loop reads no global memory.  Lane l visits partners l, l + 64, ...
// in increasing order whatever the tiling (COUL_TILE is a multiple of 64), and
// FP64 sums are combined across the wavefront, so the forces do not depend on
// the launch shape.  qtrack holds bead indices local to the replica
// (Charge_track).
#define COUL_TILE 1024
#define COUL_PER_WAVE 2
#define COUL_WAVES 4
__global__ __launch_bounds__(64 * COUL_WAVES) void k_coulomb(const RepDev* __restrict__ reps, int rep_begin,
                                                            const double* __restrict__ rx,
                                                            const double* __restrict__ ry,
                                                            const double* __restrict__ rz,
                                                            const int* __restrict__ qtrack,
                                                            const int* __restrict__ qsign, double* f) {
  __shared__ double sx[COUL_TILE], sy[COUL_TILE], sz[COUL_TILE];
  __shared__ int ss[COUL_TILE];
  const RepDev& R = reps[rep_begin + blockIdx.y];
  const int nq = R.nq;
  const int a0 = blockIdx.x * (COUL_WAVES * COUL_PER_WAVE);
  if (!R.active || a0 >= nq) return;  // uniform across the workgroup
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int ia[COUL_PER_WAVE], qa[COUL_PER_WAVE], an[COUL_PER_WAVE];
  double xa[COUL_PER_WAVE], ya[COUL_PER_WAVE], za[COUL_PER_WAVE];
  double F[COUL_PER_WAVE][3];
#pragma unroll
  for (int u = 0; u < COUL_PER_WAVE; ++u) {
    const int a = a0 + w * COUL_PER_WAVE + u;
    an[u] = a < nq ? a : -1;  // -1: no charge (never equal to a partner index)
    const int ac = a < nq ? a : 0;
    ia[u] = R.slot0 + qtrack[R.q0 + ac];
    qa[u] = qsign[R.q0 + ac];
    xa[u] = rx[ia[u]];
    ya[u] = ry[ia[u]];
    za[u] = rz[ia[u]];
    F[u][0] = F[u][1] = F[u][2] = 0.0;
  }
  for (int t0 = 0; t0 < nq; t0 += COUL_TILE) {
    const int m = min(COUL_TILE, nq - t0);
    if (t0 > 0) __syncthreads();  // the previous tile is consumed
    for (int k = threadIdx.x; k < m; k += 64 * COUL_WAVES) {
      const int ib = R.slot0 + qtrack[R.q0 + t0 + k];
      sx[k] = rx[ib];
      sy[k] = ry[ib];
      sz[k] = rz[ib];
      ss[k] = qsign[R.q0 + t0 + k];
    }
    __syncthreads();
    for (int k = lane; k < m; k += 64) {
      const int b = t0 + k;
      const double xb = sx[k], yb = sy[k], zb = sz[k];
      const int qb = ss[k];
#pragma unroll
      for (int u = 0; u < COUL_PER_WAVE; ++u) {
        if (b == an[u] || an[u] < 0) continue;
        double d[3];
        ref_nid(R, xa[u], ya[u], za[u], xb, yb, zb, d[0], d[1], d[2], false);
        double rr = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        // lambda_B qa qb e^{-kappa r} (1 + kappa r) / r^2 along d / r, with one division
        const double rd = sqrt(rr), ir = 1.0 / rd;
        const double Fc = R.lambda_b * (double)(qa[u] * qb) * exp(-R.kappa_d * rd) * (1 + rd * R.kappa_d) *
                          (ir * ir * ir);
        for (int q = 0; q < 3; ++q) F[u][q] += Fc * d[q];
      }
    }
  }
#pragma unroll
  for (int u = 0; u < COUL_PER_WAVE; ++u) {
    const double fx = warp_sum(F[u][0]), fy = warp_sum(F[u][1]), fz = warp_sum(F[u][2]);
    if (lane == 0 && an[u] >= 0) {
      f[3 * ia[u]] += fx;
      f[3 * ia[u] + 1] += fy;
      f[3 * ia[u] + 2] += fz;
    }
  }
}

void launch_coulomb(hipStream_t st, const RepDev* reps, int rep_begin, int nrep, const double* rx, const double* ry,
                    const double* rz, const int* qtrack, const int* qsign, double* f, int max_nq) {
  if (nrep <= 0 || max_nq <= 0) return;
  const int per = COUL_WAVES * COUL_PER_WAVE;
  dim3 grid((max_nq + per - 1) / per, nrep);
  k_coulomb<<<grid, 64 * COUL_WAVES, 0, st>>>(reps, rep_begin, rx, ry, rz, qtrack, qsign, f);
}
