// FP64 bead kernels: flowing-box remap (KRBC rotations + applyPBC), the
// reference's minimum image and table bins for pair batches, position update
// with the flow, displacement trackers, device noise and frame rotation of
// vectors.  The remap/min-image arithmetic is written in the reference's
// operation order with contraction off so the results match it bit for bit.
#include "common.h"

// mode 0: applyPBC only; 1: per-step remap (KRBC frame round trip, transform,
// applyPBC); 2: transformCoords; 3: inverseCoords.
__global__ void k_remap(const RepDev* __restrict__ reps, const int* __restrict__ slot_rep, int slot_begin,
                        int nslots, double* rx, double* ry, double* rz, double* dsp, int mode) {
#pragma clang fp contract(off)
  int s = slot_begin + blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= slot_begin + nslots) return;
  int r = slot_rep[s];
  if (r < 0) return;
  const RepDev& R = reps[r];
  if (!R.active || s - R.slot0 >= R.n) return;
  double x = rx[s], y = ry[s], z = rz[s];
  if (mode == 2 || mode == 3) {
    if (mode == 2) {
      double xi = x, yi = y;
      x = xi * R.ct + yi * R.st;
      y = -xi * R.st + yi * R.ct;
    } else {
      double xi = x, yi = y;
      x = xi * R.ct - yi * R.st;
      y = xi * R.st + yi * R.ct;
    }
    rx[s] = x;
    ry[s] = y;
    return;
  }
  if (mode == 1 && R.flow == 2) {
    // transformCoords; initPos; inverseCoords with the previous lattice
    // (X/main14.c:69-71), then transformCoords with the new one (X/main14.c:78).
    double x1 = x * R.ct0 + y * R.st0;
    double y1 = -x * R.st0 + y * R.ct0;
    double x2 = x1 * R.ct0 - y1 * R.st0;
    double y2 = x1 * R.st0 + y1 * R.ct0;
    x = x2 * R.ct + y2 * R.st;
    y = -x2 * R.st + y2 * R.ct;
  }
  double L2xp = R.L2xp, L2yp = R.L2yp, jump = 0.0;
  if (R.flow == 2 || !R.pbc_compat) {
    x -= L2xp * round(y / L2yp);
    y -= L2yp * round(y / L2yp);
  } else {
    // S/main26.c:1409-1428, with the +-delx shift for y crossings
    x -= L2xp * round(y / L2yp);
    if (y > L2yp / 2) {
      x -= R.delx;
      jump = -R.delx;
      y -= L2yp * round(y / L2yp);
    } else if (y < -L2yp / 2) {
      x += R.delx;
      jump = R.delx;
      y -= L2yp * round(y / L2yp);
    }
  }
  x -= R.L1xp * round((x - y * L2xp / L2yp) / R.L1xp);
  z -= R.L * round(z / R.L);
  rx[s] = x;
  ry[s] = y;
  rz[s] = z;
  if (jump != 0.0 && dsp) dsp[3 * s] += jump;
}

void launch_remap(hipStream_t st, const RepDev* reps, const int* slot_rep, int slot_begin, int nslots, double* rx,
                  double* ry, double* rz, double* dxv, double* dyv, double* dzv, double* dsp, int mode) {
  (void)dxv;
  (void)dyv;
  (void)dzv;
  if (nslots <= 0) return;
  k_remap<<<(nslots + 255) / 256, 256, 0, st>>>(reps, slot_rep, slot_begin, nslots, rx, ry, rz, dsp, mode);
}

// getNID for a batch of pairs of one replica (indices local to the replica).
__global__ void k_min_image(const RepDev* __restrict__ rep, const double* rx, const double* ry, const double* rz,
                            int npairs, const int* pi, const int* pj, double* out) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= npairs) return;
  const RepDev& R = *rep;
  int i = R.slot0 + pi[k], j = R.slot0 + pj[k];
  double ox, oy, oz;
  ref_nid(R, rx[i], ry[i], rz[i], rx[j], ry[j], rz[j], ox, oy, oz);
  out[3 * k] = ox;
  out[3 * k + 1] = oy;
  out[3 * k + 2] = oz;
}

void launch_min_image(hipStream_t st, const RepDev* rep, const double* rx, const double* ry, const double* rz,
                      int npairs, const int* pi, const int* pj, double* out) {
  if (npairs <= 0) return;
  k_min_image<<<(npairs + 255) / 256, 256, 0, st>>>(rep, rx, ry, rz, npairs, pi, pj, out);
}

// binij (S/main26.c:1518-1529): bin of getNIDbin(j,i) (S/main26.c:1458-1497,
// X/main14.c:1481-1513).  bins = {bsx,bsy,bsz}; nb = {nby,nbz,nhx,nhy,nhz,nmax}.
__global__ void k_bin_index(const RepDev* __restrict__ rep, const double* rx, const double* ry, const double* rz,
                            int npairs, const int* pi, const int* pj, const double* bins, const int* nb, int* out) {
#pragma clang fp contract(off)
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= npairs) return;
  const RepDev& R = *rep;
  int a = R.slot0 + pj[k], b = R.slot0 + pi[k];  // getNIDbin(j,i): r_j - r_i
  double dx = rx[a] - rx[b], dy = ry[a] - ry[b], dz = rz[a] - rz[b];
  int nmax = nb[5];
  dz -= R.L * round(dz / R.L);
  double dxmin = dx, dymin = dy;
  double rmin = sqrt(dx * dx + dy * dy);
  if (R.flow != 2 && R.pbc_compat) {
    if (dy > R.L2yp / 2)
      dx -= R.delx;
    else if (dy < -R.L2yp / 2)
      dx += R.delx;
  }
  for (int nx = -nmax; nx < nmax + 1; nx++)
    for (int ny = -nmax; ny < nmax + 1; ny++) {
      double dxp = dx + nx * R.L1xp + ny * R.L2xp;
      double dyp = dy + ny * R.L2yp;
      double rp = sqrt(dxp * dxp + dyp * dyp);
      if (rp < rmin) {
        dxmin = dxp;
        dymin = dyp;
        rmin = rp;
      }
    }
  double ox = dxmin, oy = dymin;
  if (R.flow == 2) {
    ox = dxmin * R.ct - dymin * R.st;
    oy = dxmin * R.st + dymin * R.ct;
  }
  int bx = (int)round(ox / bins[0]) + nb[2];
  int by = (int)round(oy / bins[1]) + nb[3];
  int bz = (int)round(dz / bins[2]) + nb[4];
  out[k] = nb[0] * nb[1] * bx + nb[1] * by + bz;
}

void launch_bin_index(hipStream_t st, const RepDev* rep, const double* rx, const double* ry, const double* rz,
                      int npairs, const int* pi, const int* pj, const double* bins, const int* nbins, int* out) {
  if (npairs <= 0) return;
  k_bin_index<<<(npairs + 255) / 256, 256, 0, st>>>(rep, rx, ry, rz, npairs, pi, pj, bins, nbins, out);
}

// Position update of updateHI (S/main26.c:841-855, X/main14.c:879-893):
// d = dt*M.F + scale*b in the wrap frame; KRBC returns positions and d to the
// lab frame (inverseCoords) before adding the flow (edot x, -edot y); shear adds
// (edot y, 0, 0).  dxv is the reference's Verlet tracker (flow included); dsp
// tracks the displacement relative to the affine flow for the mobility list.
__global__ void k_update(const RepDev* __restrict__ reps, const int* __restrict__ slot_rep, int slot_begin,
                         int nslots, double* rx, double* ry, double* rz, const double* __restrict__ mf,
                         const double* __restrict__ bz, double* dxv, double* dyv, double* dzv, double* dsp,
                         int* err) {
  int s = slot_begin + blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= slot_begin + nslots) return;
  int r = slot_rep[s];
  if (r < 0) return;
  const RepDev& R = reps[r];
  if (!R.active || s - R.slot0 >= R.n) return;
  double d[3];
  for (int c = 0; c < 3; ++c) d[c] = R.dt * mf[3 * s + c] + R.sqrt2dt * bz[3 * s + c];
  if (isfinite(d[0]) && isfinite(d[1]) && isfinite(d[2])) {
    dsp[3 * s] += d[0];
    dsp[3 * s + 1] += d[1];
    dsp[3 * s + 2] += d[2];
  }
  double x = rx[s], y = ry[s], z = rz[s];
  double ux, uy;
  if (R.flow == 2) {
    double xl = x * R.ct - y * R.st, yl = x * R.st + y * R.ct;
    double dxl = d[0] * R.ct - d[1] * R.st, dyl = d[0] * R.st + d[1] * R.ct;
    x = xl;
    y = yl;
    ux = dxl + R.dt * R.edot * x;
    uy = dyl - R.dt * R.edot * y;
  } else {
    ux = d[0] + R.dt * R.edot * y;
    uy = d[1];
  }
  double uz = d[2];
  if (!isfinite(ux) || !isfinite(uy) || !isfinite(uz)) {
    // a non-finite displacement (overflowing forces, coincident beads) is not
    // applied, so every later kernel still sees finite positions; the step
    // reports SMDLT_E_NONFINITE
    atomicOr(err, 2);
    ux = uy = uz = 0.0;
  }
  dxv[s] += ux;
  dyv[s] += uy;
  dzv[s] += uz;
  rx[s] = x + ux;
  ry[s] = y + uy;
  rz[s] = z + uz;
}

void launch_update_positions(hipStream_t st, const RepDev* reps, const int* slot_rep, int slot_begin, int nslots,
                             double* rx, double* ry, double* rz, const double* mf, const double* bz, double* dxv,
                             double* dyv, double* dzv, double* dsp, int* err) {
  if (nslots <= 0) return;
  k_update<<<(nslots + 255) / 256, 256, 0, st>>>(reps, slot_rep, slot_begin, nslots, rx, ry, rz, mf, bz, dxv, dyv,
                                                 dzv, dsp, err);
}

// Per replica: out2[2r] = max |dxv|^2, out2[2r+1] = max |dsp|^2.
__global__ void k_disp_max(const RepDev* __restrict__ reps, int rep_begin, const double* dxv, const double* dyv,
                           const double* dzv, const double* dsp, double* out2) {
  int r = rep_begin + blockIdx.x;
  const RepDev& R = reps[r];
  double m1 = 0.0, m2 = 0.0;
  for (int i = threadIdx.x; i < R.n; i += blockDim.x) {
    int s = R.slot0 + i;
    double a = dxv[s] * dxv[s] + dyv[s] * dyv[s] + dzv[s] * dzv[s];
    double b = dsp[3 * s] * dsp[3 * s] + dsp[3 * s + 1] * dsp[3 * s + 1] + dsp[3 * s + 2] * dsp[3 * s + 2];
    m1 = fmax(m1, a);
    m2 = fmax(m2, b);
  }
  __shared__ double sh[2][4];
  m1 = warp_max(m1);
  m2 = warp_max(m2);
  int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) {
    sh[0][w] = m1;
    sh[1][w] = m2;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0, b = 0;
    for (int k = 0; k < (int)(blockDim.x >> 6); ++k) {
      a = fmax(a, sh[0][k]);
      b = fmax(b, sh[1][k]);
    }
    out2[2 * r] = a;
    out2[2 * r + 1] = b;
  }
}

void launch_disp_max(hipStream_t st, const RepDev* reps, int rep_begin, int nrep, const double* dxv,
                     const double* dyv, const double* dzv, const double* dsp, double* out2) {
  if (nrep <= 0) return;
  k_disp_max<<<nrep, 256, 0, st>>>(reps, rep_begin, dxv, dyv, dzv, dsp, out2);
}

// Standard normal noise, z[3s+c], from Philox4x32-10 keyed by the replica seed
// and counted by (bead, step) + Box-Muller.  Padding slots get zeros.
__global__ void k_noise(const int* __restrict__ slot_rep, const RepDev* __restrict__ reps, int slot_begin,
                        int nslots, const uint64_t* seeds, uint64_t step, double* z) {
  int s = slot_begin + blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= slot_begin + nslots) return;
  int r = slot_rep[s];
  if (r < 0 || !reps[r].active || s - reps[r].slot0 >= reps[r].n) {
    if (r >= 0 && !reps[r].active) return;
    z[3 * s] = z[3 * s + 1] = z[3 * s + 2] = 0.0;
    return;
  }
  uint64_t key = seeds[r];
  step = reps[r].step;
  uint32_t c[4] = {(uint32_t)(s - reps[r].slot0), (uint32_t)step, (uint32_t)(step >> 32), 0x51u};
  philox10(c, (uint32_t)key, (uint32_t)(key >> 32));
  double u1 = u01_53(c[0], c[1]), u2 = u01_53(c[2], c[3]);
  double rr = sqrt(-2.0 * log(u1));
  double sn, cs;
  sincospi(2.0 * u2, &sn, &cs);
  uint32_t d[4] = {(uint32_t)(s - reps[r].slot0), (uint32_t)step, (uint32_t)(step >> 32), 0x52u};
  philox10(d, (uint32_t)key, (uint32_t)(key >> 32));
  double u3 = u01_53(d[0], d[1]), u4 = u01_53(d[2], d[3]);
  z[3 * s] = rr * cs;
  z[3 * s + 1] = rr * sn;
  z[3 * s + 2] = sqrt(-2.0 * log(u3)) * cospi(2.0 * u4);
}

void launch_noise(hipStream_t st, const int* slot_rep, const RepDev* reps, int slot_begin, int nslots,
                  const uint64_t* seeds, uint64_t step, double* z) {
  if (nslots <= 0) return;
  k_noise<<<(nslots + 255) / 256, 256, 0, st>>>(slot_rep, reps, slot_begin, nslots, seeds, step, z);
}

// Rotate AoS vectors between the lab and wrap frames (KRBC replicas only).
__global__ void k_rotate_vec(const RepDev* __restrict__ reps, const int* __restrict__ slot_rep, int slot_begin,
                             int nslots, double* v, int inverse) {
  int s = slot_begin + blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= slot_begin + nslots) return;
  int r = slot_rep[s];
  if (r < 0) return;
  const RepDev& R = reps[r];
  if (!R.active || R.flow != 2 || s - R.slot0 >= R.n) return;
  double x = v[3 * s], y = v[3 * s + 1];
  if (!inverse) {
    v[3 * s] = x * R.ct + y * R.st;
    v[3 * s + 1] = -x * R.st + y * R.ct;
  } else {
    v[3 * s] = x * R.ct - y * R.st;
    v[3 * s + 1] = x * R.st + y * R.ct;
  }
}

void launch_rotate_vec(hipStream_t st, const RepDev* reps, const int* slot_rep, int slot_begin, int nslots, double* v,
                       int inverse) {
  if (nslots <= 0) return;
  k_rotate_vec<<<(nslots + 255) / 256, 256, 0, st>>>(reps, slot_rep, slot_begin, nslots, v, inverse);
}
